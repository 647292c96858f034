/*
 * smdlt_oracle.c -- CPU FP64 restatement of the reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See smdlt_oracle.h.
 * Compiled with -ffp-contract=off; expressions keep the reference's
 * evaluation order where integer outputs depend on them.
 */
#include "smdlt_oracle.h"
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IM1 2147483563LL
#define IM2 2147483399LL
#define AM (1.0 / IM1)
#define IMM1 (IM1 - 1)
#define IA1 40014LL
#define IQ1 53668LL
#define IR1 12211LL
#define IQ2 52774LL
#define NTAB 32
#define EPS 1.2e-7
#define RNMX (1.0 - EPS)
#define NDIV (1 + IMM1 / NTAB)
#define EP 0.9624

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ======================================================================== RNG */
/* Static state of ran1/gasdev made explicit: idum2=123456789, iy=0, iv={0},
 * iset=0 (S/Initialization.c:603-606, S/main26.c:516-517). */
void orc_rng_seed(orc_rng_t* s, int64_t seed) {
  memset(s, 0, sizeof(*s));
  s->idum = seed;
  s->idum2 = 123456789;
  s->iy = 0;
  s->iset = 0;
}

/* ran1 (S/Initialization.c:601-634).  The shuffle table is filled only for a
 * non-positive seed, and idum2 is never advanced (the reference leaves out the
 * update of the second generator). */
float orc_ran1(orc_rng_t* s) {
  int j;
  int64_t k;
  float temp;
  if (s->idum <= 0) {
    if (-(s->idum) < 1)
      s->idum = 1;
    else
      s->idum = -(s->idum);
    s->idum2 = s->idum;
    for (j = NTAB + 7; j >= 0; --j) {
      k = s->idum / IQ1;
      s->idum = IA1 * (s->idum - k * IQ1) - k * IR1;
      if (s->idum < 0) s->idum += IM1;
      if (j < NTAB) s->iv[j] = s->idum;
    }
    s->iy = s->iv[0];
  }
  k = s->idum / IQ1;
  s->idum = IA1 * (s->idum - k * IQ1) - k * IR1;
  if (s->idum < 0) s->idum += IM1;
  k = s->idum2 / IQ2; /* computed and discarded, as in the reference */
  (void)k;
  if (s->idum < 0) s->idum2 += IM2;
  j = (int)(s->iy / NDIV);
  s->iy = s->iv[j] - s->idum2;
  s->iv[j] = s->idum;
  if (s->iy < 1) s->iy += IMM1;
  temp = (float)(AM * (double)s->iy);
  if ((double)temp > RNMX) return (float)RNMX;
  return temp;
}

/* gasdev (S/main26.c:515-540): polar Box-Muller in float with a double
 * log/sqrt, caching the second value. */
float orc_gasdev(orc_rng_t* s) {
  float fac, rsq, v1, v2;
  if (s->idum < 0) s->iset = 0;
  if (s->iset == 0) {
    do {
      v1 = (float)(2.0 * (double)orc_ran1(s) - 1.0);
      v2 = (float)(2.0 * (double)orc_ran1(s) - 1.0);
      rsq = v1 * v1 + v2 * v2;
    } while (rsq >= 1.0 || rsq == 0.0);
    fac = (float)sqrt(-2.0 * log((double)rsq) / (double)rsq);
    s->gset = v1 * fac;
    s->iset = 1;
    return v2 * fac;
  }
  s->iset = 0;
  return s->gset;
}

void orc_ran1_fill(orc_rng_t* s, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = orc_ran1(s);
}
void orc_gasdev_fill(orc_rng_t* s, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = (double)orc_gasdev(s);
}

/* initChains for a new run (S/main26.c:346-470), in its own control flow: flags
 * test/inittest and the per-chain attempt count initcount (checked after each
 * attempt, so the 1001st attempt restarts even when it fits). */
int orc_init_chains(orc_rng_t* s, int Nb, int Nc, double L, double* rx, double* ry, double* rz) {
  double box_side = L / 2, dx, dy, dz, theta, phi;
  int i, j, k, index, monindex, test, initcount, inittest = 0, rounds = 0;
  while (inittest == 0) {
    if (++rounds > 100000) return -1;
    inittest = 1;
    for (i = 0; i < Nc; ++i) {
      initcount = 0;
      test = 0;
      index = Nb * i;
      while (test == 0) {
        test = 1;
        if (Nc == 1) {
          rx[index] = 0;
          ry[index] = 0;
          rz[index] = 0;
        } else {
          rx[index] = orc_ran1(s) * L - box_side;
          ry[index] = orc_ran1(s) * L - box_side;
          rz[index] = orc_ran1(s) * L - box_side;
        }
        for (k = 0; k < index; k++) {
          dx = rx[index] - rx[k];
          dy = ry[index] - ry[k];
          dz = rz[index] - rz[k];
          dx -= L * round(dx / L);
          dy -= L * round(dy / L);
          dz -= L * round(dz / L);
          if (dx * dx + dy * dy + dz * dz < 10.0) test = 0;
        }
        initcount++;
        if (initcount > 1e3) {
          inittest = 0;
          break;
        }
      }
      if (inittest == 0) break;
      for (j = 1; j < Nb; j++) {
        monindex = index + j;
        test = 0;
        while (test == 0) {
          test = 1;
          theta = orc_ran1(s) * 2.0 * M_PI;
          phi = acos(2.0 * orc_ran1(s) - 1.0);
          rx[monindex] = rx[monindex - 1] + 2.05 * sin(phi) * cos(theta);
          ry[monindex] = ry[monindex - 1] + 2.05 * sin(phi) * sin(theta);
          rz[monindex] = rz[monindex - 1] + 2.05 * cos(phi);
          rx[monindex] -= L * round(rx[monindex] / L);
          ry[monindex] -= L * round(ry[monindex] / L);
          rz[monindex] -= L * round(rz[monindex] / L);
          for (k = 0; k < monindex; k++) {
            dx = rx[monindex] - rx[k];
            dy = ry[monindex] - ry[k];
            dz = rz[monindex] - rz[k];
            dx -= L * round(dx / L);
            dy -= L * round(dy / L);
            dz -= L * round(dz / L);
            if (dx * dx + dy * dy + dz * dz < 4.0) test = 0;
          }
          initcount++;
          if (initcount > 1e3) {
            inittest = 0;
            break;
          }
        }
        if (inittest == 0) break;
      }
      if (inittest == 0) break;
    }
  }
  return 0;
}

/* The charge placement of initChains (S/main26.c:485-500).  A draw that the
   float product rounds up to N (possible for N above 2^22) is redrawn. */
void orc_init_charges(orc_rng_t* s, int N, int Ncharges, int* track) {
  int* charge = (int*)calloc((size_t)N, sizeof(int));
  for (int i = 0; i < Ncharges; i++) {
    int cnumber = orc_ran1(s) * N;
    if (cnumber < N && charge[cnumber] == 0) {
      charge[cnumber] = 1;
      track[i] = cnumber;
    } else {
      i = i - 1;
    }
  }
  free(charge);
}

/* ==================================================================== Lattice */
/* initLattice (S/Initialization.c:217-298, X/Initialization.c:236-287):
 * corner points of a square of side L rotated by 90-(45+THETA0) degrees,
 * THETA0 = 0 for shear and 31.7 for KRBC (X/Parameters.h:38). */
void orc_lattice_init(orc_lattice_t* lat, int flow, double L, double edot, double dt, int pbc_compat) {
  memset(lat, 0, sizeof(*lat));
  lat->flow = flow;
  lat->L = L;
  lat->edot = edot;
  lat->dt = dt;
  lat->pbc_compat = pbc_compat;
  if (flow == 0 || edot == 0.0)
    lat->tp = INT_MAX;
  else
    lat->tp = (int)(EP / edot / dt);
  double theta0 = (flow == 2) ? 31.7 : 0.0;
  double theta = 90.0 - (45.0 + theta0);
  theta *= M_PI / 180;
  lat->point[0][0] = -sqrt(2) / 2 * L * cos(theta);
  lat->point[0][1] = sqrt(2) / 2 * L * sin(theta);
  lat->point[1][0] = -sqrt(2) / 2 * L * sin(theta);
  lat->point[1][1] = -sqrt(2) / 2 * L * cos(theta);
  lat->point[2][0] = sqrt(2) / 2 * L * cos(theta);
  lat->point[2][1] = -sqrt(2) / 2 * L * sin(theta);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) lat->point0[i][j] = lat->point[i][j];
  for (int i = 0; i < 2; ++i) {
    lat->L1[i] = lat->point[2][i] - lat->point[1][i];
    lat->L2[i] = lat->point[0][i] - lat->point[1][i];
  }
  orc_lattice_update(lat, 0);
}

/* updateLattice: shear S/main26.c:1382-1408 (deforming brick, x += y*edot*treal*dt),
 * KRBC X/main14.c:1426-1450 (x*e^{+}, y*e^{-}, then rotate into the L1 frame). */
void orc_lattice_update(orc_lattice_t* lat, uint64_t t) {
  int i;
  int treal = (lat->tp == INT_MAX) ? 0 : (int)(t % (uint64_t)lat->tp);
  lat->treal = treal;
  double flowrate = lat->edot, dt = lat->dt, L = lat->L;
  if (lat->flow == 2) {
    for (i = 0; i < 3; ++i) {
      lat->point[i][0] = lat->point0[i][0] * exp(flowrate * treal * dt);
      lat->point[i][1] = lat->point0[i][1] * exp(-flowrate * treal * dt);
    }
    for (i = 0; i < 2; ++i) {
      lat->L1[i] = lat->point[2][i] - lat->point[1][i];
      lat->L2[i] = lat->point[0][i] - lat->point[1][i];
    }
    double theta = atan(lat->L1[1] / lat->L1[0]);
    lat->L1xp = lat->L1[0] * cos(theta) + lat->L1[1] * sin(theta);
    lat->L2xp = lat->L2[0] * cos(theta) + lat->L2[1] * sin(theta);
    lat->L1yp = -lat->L1[0] * sin(theta) + lat->L1[1] * cos(theta);
    lat->L2yp = -lat->L2[0] * sin(theta) + lat->L2[1] * cos(theta);
    lat->detL = L * (-lat->L1yp * lat->L2xp + lat->L1xp * lat->L2yp);
    lat->L2pmag = sqrt(lat->L2xp * lat->L2xp + lat->L2yp * lat->L2yp);
    lat->delx = 0.0;
  } else {
    lat->point[0][0] = lat->point0[0][0] + lat->point[0][1] * flowrate * treal * dt;
    lat->point[1][0] = lat->point0[1][0] + lat->point[1][1] * flowrate * treal * dt;
    lat->point[2][0] = lat->point0[2][0] + lat->point[2][1] * flowrate * treal * dt;
    for (i = 0; i < 2; ++i) {
      lat->L1[i] = lat->point[2][i] - lat->point[1][i];
      lat->L2[i] = lat->point[0][i] - lat->point[1][i];
    }
    lat->L1xp = lat->L1[0];
    lat->L2xp = lat->L2[0];
    lat->L1yp = lat->L1[1];
    lat->L2yp = lat->L2[1];
    lat->detL = lat->L2yp * lat->L1xp * L;
    lat->L2pmag = sqrt(lat->L2xp * lat->L2xp + lat->L2yp * lat->L2yp);
    lat->delx = flowrate * treal * dt;
  }
  lat->theta = atan(lat->L1[1] / lat->L1[0]);
  lat->ct = cos(lat->theta);
  lat->st = sin(lat->theta);
}

/* applyPBC: shear S/main26.c:1409-1428 (with the extra +-delx for y crossings,
 * kept when pbc_compat), KRBC X/main14.c:1451-1461. */
void orc_apply_pbc(const orc_lattice_t* lat, int n, double* rx, double* ry, double* rz) {
  double L1xp = lat->L1xp, L2xp = lat->L2xp, L2yp = lat->L2yp, L = lat->L;
  double delx = (lat->flow == 2 || !lat->pbc_compat) ? 0.0 : lat->delx;
  for (int i = 0; i < n; ++i) {
    if (lat->flow == 2 || !lat->pbc_compat) {
      rx[i] -= L2xp * round(ry[i] / L2yp);
      ry[i] -= L2yp * round(ry[i] / L2yp);
    } else {
      rx[i] -= L2xp * round(ry[i] / L2yp);
      if (ry[i] > L2yp / 2) {
        rx[i] -= delx;
        ry[i] -= L2yp * round(ry[i] / L2yp);
      } else if (ry[i] < -L2yp / 2) {
        rx[i] += delx;
        ry[i] -= L2yp * round(ry[i] / L2yp);
      }
    }
    rx[i] -= L1xp * round((rx[i] - ry[i] * L2xp / L2yp) / L1xp);
    rz[i] -= L * round(rz[i] / L);
  }
}

/* transformCoords / inverseCoords (X/main14.c:1514-1535) */
void orc_rotate(const orc_lattice_t* lat, int n, double* rx, double* ry, int inverse) {
  double c = lat->ct, s = lat->st;
  for (int i = 0; i < n; ++i) {
    double rxi = rx[i], ryi = ry[i];
    if (!inverse) {
      rx[i] = rxi * c + ryi * s;
      ry[i] = -rxi * s + ryi * c;
    } else {
      rx[i] = rxi * c - ryi * s;
      ry[i] = rxi * s + ryi * c;
    }
  }
}

/* getNID: shear S/main26.c:1429-1457, KRBC X/main14.c:1462-1480 (vector r_i - r_j) */
void orc_nid(const orc_lattice_t* lat, const double* rx, const double* ry, const double* rz, int i,
             int j, double out[3]) {
  double L1xp = lat->L1xp, L2xp = lat->L2xp, L2yp = lat->L2yp, L = lat->L;
  double dx = rx[i] - rx[j], dy = ry[i] - ry[j], dz = rz[i] - rz[j];
  if (lat->flow == 2 || !lat->pbc_compat) {
    dx -= L2xp * round(dy / L2yp);
    dy -= L2yp * round(dy / L2yp);
  } else {
    double delx = lat->delx;
    dx -= L2xp * round(dy / L2yp);
    if (dy > L2yp / 2) {
      dx -= delx;
      dy -= L2yp * round(dy / L2yp);
    } else if (dy < -L2yp / 2) {
      dx += delx;
      dy -= L2yp * round(dy / L2yp);
    }
  }
  dx -= L1xp * round(dx / L1xp);
  dz -= L * round(dz / L);
  if (lat->flow == 2) {
    out[0] = dx * lat->ct - dy * lat->st;
    out[1] = dx * lat->st + dy * lat->ct;
  } else {
    out[0] = dx;
    out[1] = dy;
  }
  out[2] = dz;
}

/* getNIDbin: in-plane minimum over (nx,ny) in [-nmax,nmax]^2
 * (S/main26.c:1458-1497, X/main14.c:1481-1513) */
void orc_nid_bin(const orc_lattice_t* lat, const double* rx, const double* ry, const double* rz, int i,
                 int j, double out[3]) {
  int nx, ny, nmax;
  double dx, dy, dz, dxp, dyp, dxmin, dymin, rp, rmin;
  dx = rx[i] - rx[j];
  dy = ry[i] - ry[j];
  dz = rz[i] - rz[j];
  int treal = lat->treal;
  nmax = (treal <= lat->tp / 2) ? 1 : 2;
  dz -= lat->L * round(dz / lat->L);
  dxmin = dx;
  dymin = dy;
  rmin = sqrt(dx * dx + dy * dy);
  if (lat->flow != 2 && lat->pbc_compat) {
    if (dy > lat->L2yp / 2)
      dx -= lat->delx;
    else if (dy < -lat->L2yp / 2)
      dx += lat->delx;
  }
  for (nx = -nmax; nx < nmax + 1; nx++) {
    for (ny = -nmax; ny < nmax + 1; ny++) {
      dxp = dx + nx * lat->L1xp + ny * lat->L2xp;
      dyp = dy + ny * lat->L2yp;
      rp = sqrt(dxp * dxp + dyp * dyp);
      if (rp < rmin) {
        dxmin = dxp;
        dymin = dyp;
        rmin = rp;
      }
    }
  }
  if (lat->flow == 2) {
    out[0] = dxmin * lat->ct - dymin * lat->st;
    out[1] = dxmin * lat->st + dymin * lat->ct;
  } else {
    out[0] = dxmin;
    out[1] = dymin;
  }
  out[2] = dz;
}

/* Bin geometry of the far-field table (S/Initialization.c:247-296, X/Initialization.c:258-285). */
void orc_bins_init(orc_bins_t* b, const orc_lattice_t* lat, double bin_size) {
  double xmax, ymax, zmax;
  double L = lat->L;
  int tp = lat->tp;
  if (lat->flow == 2) {
    ymax = -lat->point0[1][1];
    xmax = lat->point0[2][0] * exp(lat->edot * tp * lat->dt);
  } else {
    ymax = L / 2;
    int treal = (tp == INT_MAX) ? 0 : tp;
    xmax = L / 2 * lat->edot * treal * lat->dt;
  }
  zmax = L / 2;
  b->bin_size_x = bin_size;
  b->bin_size_y = bin_size;
  b->bin_size_z = bin_size;
  b->nbx = 2 * (int)ceil(xmax / b->bin_size_x) + 1;
  b->bin_size_x = xmax / ((b->nbx - 1) / 2);
  b->nhx = (b->nbx - 1) / 2;
  b->nby = 2 * (int)ceil(ymax / b->bin_size_y) + 1;
  b->bin_size_y = ymax / ((b->nby - 1) / 2);
  b->nhy = (b->nby - 1) / 2;
  b->nbz = 2 * (int)ceil(zmax / b->bin_size_z) + 1;
  b->bin_size_z = zmax / ((b->nbz - 1) / 2);
  b->nhz = (b->nbz - 1) / 2;
  b->num_bins = b->nbx * b->nby * b->nbz;
}

/* binij (S/main26.c:1518-1529): bin of getNIDbin(j,i) */
int orc_binij(const orc_bins_t* b, const orc_lattice_t* lat, const double* rx, const double* ry,
              const double* rz, int i, int j) {
  double d[3];
  orc_nid_bin(lat, rx, ry, rz, j, i, d);
  int bx = (int)round(d[0] / b->bin_size_x) + b->nhx;
  int by = (int)round(d[1] / b->bin_size_y) + b->nhy;
  int bz = (int)round(d[2] / b->bin_size_z) + b->nhz;
  return b->nby * b->nbz * bx + b->nbz * by + bz;
}

/* ============================================================== Neighbour list */
/* verletlist (S/main26.c:682-707): for j>i, r = sqrt(|getNID(i,j)|^2) < rv, ascending j. */
void orc_verlet(const orc_lattice_t* lat, int n, const double* rx, const double* ry, const double* rz,
                double rv, int* nlist, int* list, int stride) {
  for (int i = 0; i < n; ++i) nlist[i] = 0;
  for (int i = 0; i < n - 1; ++i) {
    for (int j = i + 1; j < n; ++j) {
      double d[3];
      orc_nid(lat, rx, ry, rz, i, j, d);
      double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (r < rv) list[i * stride + nlist[i]++] = j;
    }
  }
}

/* checkVerlet (S/main26.c:858-869) */
int orc_check_verlet(int n, const double* dxv, const double* dyv, const double* dzv, double rnew) {
  for (int i = 0; i < n; ++i) {
    double d = sqrt(dxv[i] * dxv[i] + dyv[i] * dyv[i] + dzv[i] * dzv[i]);
    if (d > rnew) return 1;
  }
  return 0;
}

/* ====================================================================== Forces */
/* bondforce (S/main26.c:541-618): FENE or Hookean springs, optional bending.
 * Returns -1 on FENE overstretch (the reference prints and exits). */
int orc_bond_force(const orc_lattice_t* lat, int nb, int nc, const double* rx, const double* ry,
                   const double* rz, int spring_fene, double kappas, double qmax, double kappab,
                   double* fx, double* fy, double* fz, double tau[3]) {
  double qmax2 = qmax * qmax;
  for (int i = 0; i < nc; ++i) {
    for (int j = 1; j < nb; j++) {
      double d[3], rr, r, Fs;
      int a = nb * i + j, b = nb * i + j - 1;
      orc_nid(lat, rx, ry, rz, a, b, d);
      rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      r = sqrt(rr);
      if (spring_fene) {
        if (rr > qmax2) return -1;
        Fs = -kappas / (1.0 - rr / qmax2);
      } else {
        Fs = -kappas * (r - qmax) / r;
      }
      fx[a] += Fs * d[0];
      fy[a] += Fs * d[1];
      fz[a] += Fs * d[2];
      fx[b] += -Fs * d[0];
      fy[b] += -Fs * d[1];
      fz[b] += -Fs * d[2];
      if (tau) {
        tau[0] -= 2.0 * d[0] * Fs * d[0];
        tau[1] -= 2.0 * d[1] * Fs * d[1];
        tau[2] -= 2.0 * d[0] * Fs * d[1];
      }
    }
  }
  if (kappab > 0.0) {
    for (int i = 0; i < nc; ++i) {
      for (int j = 1; j < nb - 1; j++) {
        int c0 = nb * i + j;
        double n1[3], n2[3];
        orc_nid(lat, rx, ry, rz, c0, c0 - 1, n1);
        orc_nid(lat, rx, ry, rz, c0, c0 + 1, n2);
        double a1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
        double a2 = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
        double var = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (a1 * a2);
        double thetab = M_PI - acos(var);
        double coeff = -kappab * thetab / sqrt(1 - var * var);
        double* f[3] = {fx, fy, fz};
        for (int c = 0; c < 3; ++c) {
          f[c][c0] += coeff * ((n1[c] + n2[c]) / (a1 * a2) - var * (n1[c] / (a1 * a1) + n2[c] / (a2 * a2)));
          f[c][c0 - 1] += coeff * (var * n1[c] / (a1 * a1) - n2[c] / (a1 * a2));
          f[c][c0 + 1] += coeff * (var * n2[c] / (a2 * a2) - n1[c] / (a1 * a2));
        }
        if (tau) {  /* S/main26.c:606-613: xx and yy only; two yy terms read the z components */
          tau[0] += n1[0] * coeff * (var * n1[0] / (a1 * a1) - n2[0] / (a1 * a2));
          tau[0] -= n1[0] * coeff * ((n1[0]) / (a1 * a2) - var * (n1[0] / (a1 * a1)));
          tau[0] += n2[0] * coeff * ((n2[0]) / (a1 * a2) - var * (n2[0] / (a2 * a2)));
          tau[0] -= n2[0] * coeff * (var * n2[0] / (a2 * a2) - n1[0] / (a1 * a2));
          tau[1] += n1[1] * coeff * (var * n1[1] / (a1 * a1) - n2[1] / (a1 * a2));
          tau[1] -= n1[1] * coeff * ((n1[2]) / (a1 * a2) - var * (n1[2] / (a1 * a1)));
          tau[1] += n2[1] * coeff * ((n2[2]) / (a1 * a2) - var * (n2[2] / (a2 * a2)));
          tau[1] -= n2[1] * coeff * (var * n2[1] / (a2 * a2) - n1[1] / (a1 * a2));
        }
      }
    }
  }
  return 0;
}

/* Chain observables of the output step (S/main26.c:74-93 unwrapping, 925-962
 * calcExt/calcRg, 2391-2415 calcCOM): p_{j+1} = p_j + getNID(j, j+1) from
 * p_0 = 0, so each chain is unwrapped (mirrored: getNID is r_j - r_{j+1});
 * ext = max - min of p_x; com = mean of p; rg = mean of (p - com)(p - com) as
 * xx, xy, xz, yy, yz, zz.  The reference's calcCOM and calcRg accumulators are
 * neither initialised nor reset between chains (undefined behaviour); this is
 * the per-chain value they are meant to hold. */
void orc_chain_obs(const orc_lattice_t* lat, int nb, int nc, const double* rx, const double* ry,
                   const double* rz, double* ext, double* com, double* rg) {
  double* p = (double*)malloc(sizeof(double) * 3 * (size_t)nb);
  for (int i = 0; i < nc; ++i) {
    p[0] = p[1] = p[2] = 0.0;
    for (int j = 0; j < nb - 1; ++j) {
      double d[3];
      orc_nid(lat, rx, ry, rz, i * nb + j, i * nb + j + 1, d);
      for (int k = 0; k < 3; ++k) p[3 * (j + 1) + k] = p[3 * j + k] + d[k];
    }
    double mx = p[0], mn = p[0], c[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < nb; ++j) {
      mx = fmax(mx, p[3 * j]);
      mn = fmin(mn, p[3 * j]);
      for (int k = 0; k < 3; ++k) c[k] += p[3 * j + k];
    }
    for (int k = 0; k < 3; ++k) c[k] /= nb;
    double g[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nb; ++j) {
      double a = p[3 * j] - c[0], b = p[3 * j + 1] - c[1], e = p[3 * j + 2] - c[2];
      g[0] += a * a;
      g[1] += a * b;
      g[2] += a * e;
      g[3] += b * b;
      g[4] += b * e;
      g[5] += e * e;
    }
    ext[i] = mx - mn;
    for (int k = 0; k < 3; ++k) com[3 * i + k] = c[k];
    for (int k = 0; k < 6; ++k) rg[6 * i + k] = g[k] / nb;
  }
  free(p);
}

/* LJforce over the Verlet list: shear 12eps(r6^2-r6)/r^2 (S/main26.c:619-681),
 * extensional 48eps(r6^2-r6/2)/r^2 (X/main14.c:659), (sigma/r)^6 = (4/r^2)^3. */
void orc_lj_force(const orc_lattice_t* lat, int n, const double* rx, const double* ry, const double* rz,
                  const int* nlist, const int* list, int stride, double r2cut, double epsilon,
                  int ext_form, double* fx, double* fy, double* fz, double tau[3]) {
  for (int i = 0; i < n; ++i) {
    for (int jj = 0; jj < nlist[i]; ++jj) {
      int k = list[i * stride + jj];
      double d[3];
      orc_nid(lat, rx, ry, rz, i, k, d);
      double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (rr < r2cut) {
        double ratio = 4.00 / rr;
        double r6 = ratio * ratio * ratio;
        double coeff = ext_form ? (48.0 * epsilon / rr) * (r6 * r6 - 0.5 * r6)
                                : (12 * epsilon / rr) * (r6 * r6 - r6);
        fx[i] += coeff * d[0];
        fy[i] += coeff * d[1];
        fz[i] += coeff * d[2];
        fx[k] -= coeff * d[0];
        fy[k] -= coeff * d[1];
        fz[k] -= coeff * d[2];
        if (tau) {
          tau[0] -= 2.0 * d[0] * coeff * d[0];
          tau[1] -= 2.0 * d[1] * coeff * d[1];
          tau[2] -= 2.0 * d[0] * coeff * d[1];
        }
      }
    }
  }
}

/* coulombForce (S/main26.c:1891-1925): all charged pairs, no cutoff.  q may be
 * any integer charge; the reference uses Charge in {0,1}. */
void orc_coulomb_force(const orc_lattice_t* lat, int n, const double* rx, const double* ry,
                       const double* rz, const int* q, double lambda_b, double lambda_d, double* fx,
                       double* fy, double* fz) {
  double kappa_d = 1 / lambda_d;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      if (q[i] != 0 && q[j] != 0) {
        double d[3];
        orc_nid(lat, rx, ry, rz, i, j, d);
        double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double r = sqrt(rr);
        double Fc = lambda_b * q[i] * q[j] * exp(-kappa_d * r) * (1 + r * kappa_d) / rr;
        fx[i] += Fc * d[0] / r;
        fy[i] += Fc * d[1] / r;
        fz[i] += Fc * d[2] / r;
        fx[j] -= Fc * d[0] / r;
        fy[j] -= Fc * d[1] / r;
        fz[j] -= Fc * d[2] / r;
      }
    }
  }
}

/* screened-Coulomb energy, the double loop of chargeHop (S/main26.c:1955-1970) */
double orc_coulomb_energy(const orc_lattice_t* lat, int n, const double* rx, const double* ry,
                          const double* rz, const int* q, double lambda_b, double lambda_d) {
  double kappa_d = 1 / lambda_d, E = 0;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      if (q[i] != 0 && q[j] != 0) {
        double d[3];
        orc_nid(lat, rx, ry, rz, i, j, d);
        double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double r = sqrt(rr);
        E += (lambda_b * q[i] * q[j] / r) * exp(-kappa_d * r);
      }
    }
  }
  return E;
}

/* ============================================================ Ewald-summed RPY */
double orc_ewald_selfconst(double alpha) {
  double rtpi = sqrt(M_PI);
  double alpha3 = alpha * alpha * alpha;
  return 1 - 6 / rtpi * alpha + 40 / (3 * rtpi) * alpha3;
}

/* Beenakker real-space coefficients (S/main26.c:1727-1745): M1 = t1 I + t2 dd^T/r^2.
 * overlap_mode 0 adds the exact overlap correction RPY_ov - RPY_far for r<2;
 * mode 1 uses the reference's r<=2 branch. */
void orc_rpy_real_coeffs(double r, double alpha, int overlap_mode, double* t1, double* t2) {
  double rr = r * r, rtpi = sqrt(M_PI);
  double a2 = alpha * alpha, a3 = a2 * alpha, a4 = a2 * a2, a5 = a4 * alpha, a7 = a4 * a2 * alpha;
  double C1, C2, C3, C4;
  if (overlap_mode == 1 && !(r > 2.0)) {
    C1 = 1.0 - 0.28125 * r;
    C2 = 64.0 * a7 + 12.0 * a3 - 80.0 * a5 - 4.5 * alpha + 14.0 * a3 + 0.25 * alpha;
    C3 = 0.09375 * r;
    C4 = -64.0 * a7 - 12.0 * a3 + 64.0 * a5 + 1.5 * alpha - 2.0 * a3 - 0.75 * alpha;
  } else {
    C1 = 0.75 / r + 0.5 / (r * rr);
    C2 = 4.0 * a7 * rr * rr + 3.0 * a3 * rr - 20.0 * a5 * rr - 4.5 * alpha + 14.0 * a3 + alpha / rr;
    C3 = 0.75 / r - 1.5 / (r * rr);
    C4 = -4.0 * a7 * rr * rr - 3.0 * a3 * rr + 16.0 * a5 * rr + 1.5 * alpha - 2.0 * a3 - 3.0 * alpha / rr;
  }
  double ec = erfc(alpha * r), g = exp(-a2 * rr) / rtpi;
  *t1 = C1 * ec + C2 * g;
  *t2 = C3 * ec + C4 * g;
  if (overlap_mode == 0 && r < 2.0) {
    *t1 += (1.0 - 9.0 * r / 32.0) - (0.75 / r + 0.5 / (r * rr));
    *t2 += (3.0 * r / 32.0) - (0.75 / r - 1.5 / (r * rr));
  }
}

static void rpy_free_block(const double d[3], double B[9]) {
  double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(rr);
  double t1, t2;
  if (r >= 2.0) {
    t1 = 0.75 / r + 0.5 / (r * rr);
    t2 = 0.75 / r - 1.5 / (r * rr);
  } else {
    t1 = 1.0 - 9.0 * r / 32.0;
    t2 = 3.0 * r / 32.0;
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) B[3 * a + b] = (a == b ? t1 : 0.0) + t2 * d[a] * d[b] / rr;
}

void orc_rpy_free_dense(int n, const double* pos, double* M) {
  int n3 = 3 * n;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      double B[9] = {0};
      if (i == j) {
        B[0] = B[4] = B[8] = 1.0;
      } else {
        double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
        rpy_free_block(d, B);
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) M[(size_t)(3 * i + a) * n3 + 3 * j + b] = B[3 * a + b];
    }
  }
}

typedef struct {
  double k[3];
  double m2;
} kvec_t;

/* Reciprocal vectors k = m1 b1 + m2 b2 + m3 b3 with a_i . b_j = 2 pi delta_ij,
 * 0 < |k| < kcut; m2(k) of S/main26.c:1622 without its half-space factor 2. */
static kvec_t* build_kvecs(const orc_cell_t* c, double alpha, double kcut, int* count) {
  double area = c->a1x * c->a2y - c->a1y * c->a2x;
  double b1[2] = {2 * M_PI * c->a2y / area, -2 * M_PI * c->a2x / area};
  double b2[2] = {-2 * M_PI * c->a1y / area, 2 * M_PI * c->a1x / area};
  double b3 = 2 * M_PI / c->Lz;
  double la1 = sqrt(c->a1x * c->a1x + c->a1y * c->a1y), la2 = sqrt(c->a2x * c->a2x + c->a2y * c->a2y);
  int m1max = (int)ceil(kcut * la1 / (2 * M_PI)) + 1;
  int m2max = (int)ceil(kcut * la2 / (2 * M_PI)) + 1;
  int m3max = (int)ceil(kcut * c->Lz / (2 * M_PI)) + 1;
  int cap = (2 * m1max + 1) * (2 * m2max + 1) * (2 * m3max + 1);
  kvec_t* kv = (kvec_t*)malloc(sizeof(kvec_t) * (size_t)cap);
  int cnt = 0;
  double a2 = alpha * alpha, a4 = a2 * a2;
  for (int m1 = -m1max; m1 <= m1max; ++m1)
    for (int m2 = -m2max; m2 <= m2max; ++m2)
      for (int m3 = -m3max; m3 <= m3max; ++m3) {
        if (!m1 && !m2 && !m3) continue;
        double kx = m1 * b1[0] + m2 * b2[0], ky = m1 * b1[1] + m2 * b2[1], kz = m3 * b3;
        double kk = kx * kx + ky * ky + kz * kz;
        if (kk >= kcut * kcut) continue;
        kv[cnt].k[0] = kx;
        kv[cnt].k[1] = ky;
        kv[cnt].k[2] = kz;
        kv[cnt].m2 = (1.0 - kk / 3.0) * (1.0 + kk / (4.0 * a2) + kk * kk / (8.0 * a4)) * 6.0 * M_PI / kk *
                     exp(-kk / (4.0 * a2));
        cnt++;
      }
  *count = cnt;
  return kv;
}

void orc_rpy_ewald_dense(const orc_cell_t* c, int n, const double* pos, double alpha, double rcut,
                         double kcut, int overlap_mode, double* M) {
  int n3 = 3 * n, nk = 0;
  double area = c->a1x * c->a2y - c->a1y * c->a2x;
  double V = fabs(area) * c->Lz;
  kvec_t* kv = build_kvecs(c, alpha, kcut, &nk);
  double la1 = sqrt(c->a1x * c->a1x + c->a1y * c->a1y), la2 = sqrt(c->a2x * c->a2x + c->a2y * c->a2y);
  double w1 = fabs(area) / la2, w2 = fabs(area) / la1; /* widths across a1 and a2 */
  int n1max = (int)ceil(rcut / w1) + 1, n2max = (int)ceil(rcut / w2) + 1, n3max = (int)ceil(rcut / c->Lz) + 1;
  double self = orc_ewald_selfconst(alpha);
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; ++i) {
    for (int j = i; j < n; ++j) {
      double B[9] = {0};
      double d0[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
      for (int p1 = -n1max; p1 <= n1max; ++p1)
        for (int p2 = -n2max; p2 <= n2max; ++p2)
          for (int p3 = -n3max; p3 <= n3max; ++p3) {
            if (i == j && !p1 && !p2 && !p3) continue;
            double d[3] = {d0[0] + p1 * c->a1x + p2 * c->a2x, d0[1] + p1 * c->a1y + p2 * c->a2y, d0[2] + p3 * c->Lz};
            double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(rr);
            if (r >= rcut) continue;
            double t1, t2;
            orc_rpy_real_coeffs(r, alpha, overlap_mode, &t1, &t2);
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b) B[3 * a + b] += (a == b ? t1 : 0.0) + t2 * d[a] * d[b] / rr;
          }
      for (int q = 0; q < nk; ++q) {
        const double* k = kv[q].k;
        double kk = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        double w = kv[q].m2 / V * cos(k[0] * d0[0] + k[1] * d0[1] + k[2] * d0[2]);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) B[3 * a + b] += w * ((a == b ? 1.0 : 0.0) - k[a] * k[b] / kk);
      }
      if (i == j) {
        B[0] += self;
        B[4] += self;
        B[8] += self;
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          M[(size_t)(3 * i + a) * n3 + 3 * j + b] = B[3 * a + b];
          M[(size_t)(3 * j + b) * n3 + 3 * i + a] = B[3 * a + b];
        }
    }
  }
  free(kv);
}

/* u = M v for the Ewald-summed RPY matrix of orc_rpy_ewald_dense without forming
 * M, so that systems of 10^4-10^5 beads can be checked.  Same split and terms,
 * arranged differently:
 *  - real space: the pair vector is wrapped to fractional coordinates in
 *    [-1/2, 1/2] of a Gauss-reduced basis; images p with |s_i + p_i| below
 *    rcut/w_i + 1/2 (w_i the width across basis vector i) are the only ones that
 *    can lie within rcut, so p_i runs over [-m_i, m_i], m_i = floor(rcut/w_i + 1/2);
 *  - reciprocal space: cos(k.(r_j - r_i)) = Re(e^{i k.r_j} e^{-i k.r_i}), so
 *    sum_j M2(k) v_j cos(..) = Re(conj(e^{i k.r_i}) S(k)) with the structure
 *    factor S(k) = sum_j v_j e^{i k.r_j}; k and -k give equal terms, so the half
 *    space is summed twice.  e^{i k.r} is the product of per-bead phase tables
 *    e^{i m_d b_d.r} (one sincos per bead and index);
 *  - the self term selfconst * v_i.
 * Deterministic: each structure factor and each u_i is summed by one thread in
 * a fixed order. */
void orc_rpy_ewald_matvec(const orc_cell_t* c0, int n, const double* pos, double alpha, double rcut, double kcut,
                          int overlap_mode, const double* v, double* u) {
  /* Gauss-reduce the in-plane basis */
  double a1[2] = {c0->a1x, c0->a1y}, a2[2] = {c0->a2x, c0->a2y};
  for (int it = 0; it < 100; ++it) {
    double n1 = a1[0] * a1[0] + a1[1] * a1[1], n2 = a2[0] * a2[0] + a2[1] * a2[1];
    if (n2 < n1) {
      double t0 = a1[0], t1 = a1[1];
      a1[0] = a2[0];
      a1[1] = a2[1];
      a2[0] = t0;
      a2[1] = t1;
      n1 = n2;
    }
    double mu = round((a1[0] * a2[0] + a1[1] * a2[1]) / n1);
    if (mu == 0.0) break;
    a2[0] -= mu * a1[0];
    a2[1] -= mu * a1[1];
  }
  const double Lz = c0->Lz;
  const double area = a1[0] * a2[1] - a1[1] * a2[0], aarea = fabs(area), V = aarea * Lz;
  const double la1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1]), la2 = sqrt(a2[0] * a2[0] + a2[1] * a2[1]);
  const double w[3] = {aarea / la2, aarea / la1, Lz};
  int m[3];
  for (int d = 0; d < 3; ++d) m[d] = (int)floor(rcut / w[d] + 0.5);
  const double self = orc_ewald_selfconst(alpha);
  /* real space and self term */
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    double acc[3] = {self * v[3 * i], self * v[3 * i + 1], self * v[3 * i + 2]};
    for (int j = 0; j < n; ++j) {
      double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1], dz = pos[3 * j + 2] - pos[3 * i + 2];
      double s1 = (dx * a2[1] - dy * a2[0]) / area, s2 = (a1[0] * dy - a1[1] * dx) / area, s3 = dz / Lz;
      s1 -= round(s1);
      s2 -= round(s2);
      s3 -= round(s3);
      for (int p1 = -m[0]; p1 <= m[0]; ++p1)
        for (int p2 = -m[1]; p2 <= m[1]; ++p2)
          for (int p3 = -m[2]; p3 <= m[2]; ++p3) {
            double f1 = s1 + p1, f2 = s2 + p2, f3 = s3 + p3;
            if (i == j && f1 == 0.0 && f2 == 0.0 && f3 == 0.0) continue;
            double d[3] = {f1 * a1[0] + f2 * a2[0], f1 * a1[1] + f2 * a2[1], f3 * Lz};
            double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (rr >= rcut * rcut) continue;
            double r = sqrt(rr), t1, t2;
            orc_rpy_real_coeffs(r, alpha, overlap_mode, &t1, &t2);
            double dv = (d[0] * v[3 * j] + d[1] * v[3 * j + 1] + d[2] * v[3 * j + 2]) * t2 / rr;
            for (int a = 0; a < 3; ++a) acc[a] += t1 * v[3 * j + a] + dv * d[a];
          }
    }
    for (int a = 0; a < 3; ++a) u[3 * i + a] = acc[a];
  }
  /* reciprocal basis b_i . a_j = 2 pi delta_ij */
  const double b1[2] = {2 * M_PI * a2[1] / area, -2 * M_PI * a2[0] / area};
  const double b2[2] = {-2 * M_PI * a1[1] / area, 2 * M_PI * a1[0] / area};
  const double b3 = 2 * M_PI / Lz;
  const int M1 = (int)ceil(kcut * la1 / (2 * M_PI)) + 1, M2 = (int)ceil(kcut * la2 / (2 * M_PI)) + 1,
            M3 = (int)ceil(kcut * Lz / (2 * M_PI)) + 1;
  /* half space: m1 > 0, or m1 = 0 and m2 > 0, or m1 = m2 = 0 and m3 > 0 */
  int cap = (M1 + 1) * (2 * M2 + 1) * (2 * M3 + 1), nk = 0;
  int* km = (int*)malloc(sizeof(int) * 3 * (size_t)cap);
  double* kw = (double*)malloc(sizeof(double) * 4 * (size_t)cap);
  const double a2s = alpha * alpha, a4s = a2s * a2s;
  for (int q1 = 0; q1 <= M1; ++q1)
    for (int q2 = -M2; q2 <= M2; ++q2)
      for (int q3 = -M3; q3 <= M3; ++q3) {
        if (q1 == 0 && (q2 < 0 || (q2 == 0 && q3 <= 0))) continue;
        double kx = q1 * b1[0] + q2 * b2[0], ky = q1 * b1[1] + q2 * b2[1], kz = q3 * b3;
        double kk = kx * kx + ky * ky + kz * kz;
        if (kk >= kcut * kcut) continue;
        km[3 * nk] = q1;
        km[3 * nk + 1] = q2;
        km[3 * nk + 2] = q3;
        kw[4 * nk] = kx;
        kw[4 * nk + 1] = ky;
        kw[4 * nk + 2] = kz;
        kw[4 * nk + 3] = 2.0 * (1.0 - kk / 3.0) * (1.0 + kk / (4.0 * a2s) + kk * kk / (8.0 * a4s)) * 6.0 * M_PI / kk *
                         exp(-kk / (4.0 * a2s)) / V;
        nk++;
      }
  /* phase tables [index][bead] of e^{i m b_d . r} */
  const int n1t = M1 + 1, n2t = 2 * M2 + 1, n3t = 2 * M3 + 1;
  double* T1 = (double*)malloc(sizeof(double) * 2 * (size_t)n1t * n);
  double* T2 = (double*)malloc(sizeof(double) * 2 * (size_t)n2t * n);
  double* T3 = (double*)malloc(sizeof(double) * 2 * (size_t)n3t * n);
#pragma omp parallel for
  for (int j = 0; j < n; ++j) {
    double p1 = b1[0] * pos[3 * j] + b1[1] * pos[3 * j + 1], p2 = b2[0] * pos[3 * j] + b2[1] * pos[3 * j + 1],
           p3 = b3 * pos[3 * j + 2];
    for (int q = 0; q < n1t; ++q) {
      T1[2 * ((size_t)q * n + j)] = cos(q * p1);
      T1[2 * ((size_t)q * n + j) + 1] = sin(q * p1);
    }
    for (int q = -M2; q <= M2; ++q) {
      T2[2 * ((size_t)(q + M2) * n + j)] = cos(q * p2);
      T2[2 * ((size_t)(q + M2) * n + j) + 1] = sin(q * p2);
    }
    for (int q = -M3; q <= M3; ++q) {
      T3[2 * ((size_t)(q + M3) * n + j)] = cos(q * p3);
      T3[2 * ((size_t)(q + M3) * n + j) + 1] = sin(q * p3);
    }
  }
  /* projected, weighted structure factors S'(k) = w(k) (I - k k^T/k^2) S(k) */
  double* Sk = (double*)malloc(sizeof(double) * 6 * (size_t)(nk ? nk : 1));
#pragma omp parallel for schedule(dynamic, 4)
  for (int q = 0; q < nk; ++q) {
    const double* t1 = T1 + 2 * (size_t)km[3 * q] * n;
    const double* t2 = T2 + 2 * (size_t)(km[3 * q + 1] + M2) * n;
    const double* t3 = T3 + 2 * (size_t)(km[3 * q + 2] + M3) * n;
    double sr[3] = {0, 0, 0}, si[3] = {0, 0, 0};
    for (int j = 0; j < n; ++j) {
      double er = t1[2 * j] * t2[2 * j] - t1[2 * j + 1] * t2[2 * j + 1];
      double ei = t1[2 * j] * t2[2 * j + 1] + t1[2 * j + 1] * t2[2 * j];
      double fr = er * t3[2 * j] - ei * t3[2 * j + 1], fi = er * t3[2 * j + 1] + ei * t3[2 * j];
      for (int a = 0; a < 3; ++a) {
        sr[a] += v[3 * j + a] * fr;
        si[a] += v[3 * j + a] * fi;
      }
    }
    const double* k = kw + 4 * q;
    double kk = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    double pr = (k[0] * sr[0] + k[1] * sr[1] + k[2] * sr[2]) / kk, pi = (k[0] * si[0] + k[1] * si[1] + k[2] * si[2]) / kk;
    for (int a = 0; a < 3; ++a) {
      Sk[6 * q + a] = k[3] * (sr[a] - pr * k[a]);
      Sk[6 * q + 3 + a] = k[3] * (si[a] - pi * k[a]);
    }
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    double acc[3] = {0, 0, 0};
    for (int q = 0; q < nk; ++q) {
      const double* t1 = T1 + 2 * ((size_t)km[3 * q] * n + i);
      const double* t2 = T2 + 2 * ((size_t)(km[3 * q + 1] + M2) * n + i);
      const double* t3 = T3 + 2 * ((size_t)(km[3 * q + 2] + M3) * n + i);
      double er = t1[0] * t2[0] - t1[1] * t2[1], ei = t1[0] * t2[1] + t1[1] * t2[0];
      double fr = er * t3[0] - ei * t3[1], fi = er * t3[1] + ei * t3[0];
      /* Re(S' conj(e)) */
      for (int a = 0; a < 3; ++a) acc[a] += Sk[6 * q + a] * fr + Sk[6 * q + 3 + a] * fi;
    }
    for (int a = 0; a < 3; ++a) u[3 * i + a] += acc[a];
  }
  free(Sk);
  free(T1);
  free(T2);
  free(T3);
  free(km);
  free(kw);
}

/* The reciprocal-space sum of calcHI and ewaldBin, restated with its range logic
 * (S/main26.c:1593-1652 and 1000-1097, X/main14.c:1034-1131): the wave vectors
 * (kx >= 0 half space in the reference's index names, shear rkx from ky and rky
 * from kx, KRBC from the deformed lattice with kcoeff = 2 pi/detL), the first/last
 * in-cut bounds along ky and kz, the while-loop that fixes kxmax, M2(k) per vector,
 * the kx = 0 plane halved and k = 0 removed.  Returns the count; rk[3q] and
 * M2[9q] are malloc'ed. */
static int ref_kspace(const orc_lattice_t* lat, double M_HI, double rce, double box_volume, double** rk_out,
                      double** M2_out) {
  int kx, ky, kz, kxmax, kymax, kzmax, ylower, zlower, yupper, zupper;
  double rkx = 0, rky = 0, rkz, rkk, rkxy, m2;
  const double* L1 = lat->L1;
  const double* L2 = lat->L2;
  const double L = lat->L, L2yp = lat->L2yp, detL = lat->detL;
  const int krbc = (lat->flow == 2);
  const double alpha = M_HI / rce, alpha2 = alpha * alpha, alpha4 = alpha2 * alpha2;
  const double rkcut = 2.0 * M_HI * M_HI / rce, rkkcut = rkcut * rkcut;
  kxmax = (int)ceil(rkcut * detL / (M_PI * 2.0 * L * L2[1]));
  kymax = (int)ceil(rkcut * detL / (M_PI * 2.0 * L * L1[0]));
  kzmax = (int)ceil(rkcut * L / (M_PI * 2.0));
  const int kxalloc = kxmax + 1;
  int* kyupper = (int*)calloc((size_t)kxalloc + 2, sizeof(int));
  int* kylower = (int*)calloc((size_t)kxalloc + 2, sizeof(int));
  int* kzupper = (int*)calloc((size_t)(kxalloc + 2) * (2 * kymax + 1), sizeof(int));
  int* kzlower = (int*)calloc((size_t)(kxalloc + 2) * (2 * kymax + 1), sizeof(int));
  const double kcoeff = krbc ? M_PI * 2.0 / detL : M_PI * 2.0;
#define RK_SET()                                             \
  do {                                                       \
    if (krbc) {                                              \
      rkx = kcoeff * L * (kx * L2[1] - ky * L1[1]);          \
      rky = kcoeff * L * (-kx * L2[0] + ky * L1[0]);         \
    } else {                                                 \
      rky = kcoeff * (kx - ky * (L2[0] / L1[0])) / L2[1];    \
      rkx = kcoeff * ky / L1[0];                             \
    }                                                        \
  } while (0)
  for (kx = 0; kx < kxmax + 1; kx++) {
    ylower = 0;
    yupper = 0;
    for (ky = -kymax; ky < kymax + 1; ky++) {
      RK_SET();
      rkxy = rkx * rkx + rky * rky;
      zlower = 0;
      zupper = 0;
      for (kz = -kzmax; kz < kzmax + 1; kz++) {
        rkz = kz * M_PI * 2.0 / L;
        rkk = rkxy + rkz * rkz;
        if (rkk < rkkcut && zlower == 0) {
          kzlower[(2 * kymax + 1) * kx + ky + kymax] = kz;
          zlower = 1;
        }
        if (rkk > rkkcut && zlower == 1 && zupper == 0) {
          kzupper[(2 * kymax + 1) * kx + ky + kymax] = kz - 1;
          zupper = 1;
        }
      }
      if (rkxy < rkkcut && ylower == 0) {
        kylower[kx] = ky;
        ylower = 1;
      }
      if (rkxy > rkkcut && ylower == 1 && yupper == 0) {
        kyupper[kx] = ky - 1;
        yupper = 1;
      }
    }
  }
  kx = 0;
  while (kx < kxalloc && kyupper[kx] != -kymax && kx <= (int)(rkcut * detL / (M_PI * 2.0 * L * L2yp))) {
    kxmax = kx;
    kx++;
  }
  kxmax++;
  if (kxmax > kxalloc - 1) kxmax = kxalloc - 1; /* stay inside the arrays the reference allocated */
  int kcount = 0;
  for (kx = 0; kx < kxmax + 1; kx++)
    for (ky = kylower[kx]; ky < kyupper[kx] + 1; ky++)
      for (kz = kzlower[(2 * kymax + 1) * kx + ky + kymax]; kz < kzupper[(2 * kymax + 1) * kx + ky + kymax] + 1; kz++)
        kcount++;
  double* M2 = (double*)calloc((size_t)9 * (kcount + 1), sizeof(double));
  double* rk = (double*)calloc((size_t)3 * (kcount + 1), sizeof(double));
  int q = 0;
  for (kx = 0; kx < kxmax + 1; kx++) {
    for (ky = kylower[kx]; ky < kyupper[kx] + 1; ky++) {
      RK_SET();
      rkxy = rkx * rkx + rky * rky;
      for (kz = kzlower[(2 * kymax + 1) * kx + ky + kymax]; kz < kzupper[(2 * kymax + 1) * kx + ky + kymax] + 1; kz++) {
        rkz = kz * M_PI * 2.0 / L;
        rkk = rkxy + rkz * rkz;
        m2 = 2.0 * (1.0 - rkk / 3.0) * (1.0 + rkk / (4.0 * alpha2) + rkk * rkk / (8.0 * alpha4)) * 6.0 * M_PI / rkk *
             exp(-rkk / (4.0 * alpha2));
        double* m = M2 + 9 * q;
        m[0] = m2 * (1 - rkx * rkx / rkk) / box_volume;
        m[1] = m2 * (-rkx * rky / rkk) / box_volume;
        m[2] = m2 * (-rkx * rkz / rkk) / box_volume;
        m[3] = m2 * (-rky * rkx / rkk) / box_volume;
        m[4] = m2 * (1 - rky * rky / rkk) / box_volume;
        m[5] = m2 * (-rky * rkz / rkk) / box_volume;
        m[6] = m2 * (-rkz * rkx / rkk) / box_volume;
        m[7] = m2 * (-rkz * rky / rkk) / box_volume;
        m[8] = m2 * (1 - rkz * rkz / rkk) / box_volume;
        rk[3 * q] = rkx;
        rk[3 * q + 1] = rky;
        rk[3 * q + 2] = rkz;
        q++;
      }
    }
  }
  q = 0;
  for (ky = kylower[0]; ky < kyupper[0] + 1; ky++) {
    for (kz = kzlower[ky + kymax]; kz < kzupper[ky + kymax] + 1; kz++) {
      for (int i = 0; i < 9; i++) M2[9 * q + i] /= 2.0;
      if (ky == 0 && kz == 0)
        for (int i = 0; i < 9; i++) M2[9 * q + i] = 0.0;
      q++;
    }
  }
#undef RK_SET
  free(kyupper);
  free(kylower);
  free(kzupper);
  free(kzlower);
  *rk_out = rk;
  *M2_out = M2;
  return kcount;
}

/* The self block of calcHI and ewaldBin (S/main26.c:1653-1686, 1097-1130): selfconst,
 * the reciprocal diagonal and the image sum with its quirks (only images with all
 * three indices nonzero; the z entry adds dy*dy; no r < rce test). */
static void ref_self(const orc_lattice_t* lat, double alpha, int nmax, int kcount, const double* M2, double Dtr[3]) {
  const double* L1 = lat->L1;
  const double* L2 = lat->L2;
  const double L = lat->L, rtpi = sqrt(M_PI);
  const double alpha2 = alpha * alpha, alpha3 = alpha2 * alpha, alpha4 = alpha2 * alpha2, alpha5 = alpha4 * alpha,
               alpha7 = alpha4 * alpha2 * alpha;
  const double selfconst = 1 - 6 / rtpi * alpha + 40 / (3 * rtpi) * alpha3;
  Dtr[0] = selfconst;
  Dtr[1] = selfconst;
  Dtr[2] = selfconst;
  for (int q = 0; q < kcount; ++q) {
    Dtr[0] += M2[9 * q];
    Dtr[1] += M2[9 * q + 4];
    Dtr[2] += M2[9 * q + 8];
  }
  for (int nx = -nmax; nx < nmax + 1; nx++)
    for (int ny = -nmax; ny < nmax + 1; ny++)
      for (int nz = -nmax; nz < nmax + 1; nz++) {
        if (nx != 0 && ny != 0 && nz != 0) {
          double dx = nx * L1[0] + ny * L2[0], dy = nx * L1[1] + ny * L2[1], dz = nz * L;
          double rr = dx * dx + dy * dy + dz * dz, r = sqrt(rr);
          double C1 = 0.75 / r + 0.5 / (r * rr);
          double C2 = 4.0 * alpha7 * rr * rr + 3.0 * alpha3 * rr - 20.0 * alpha5 * rr - 4.5 * alpha + 14.0 * alpha3 + alpha / rr;
          double C3 = 0.75 / r - 1.5 / (r * rr);
          double C4 = -4.0 * alpha7 * rr * rr - 3.0 * alpha3 * rr + 16.0 * alpha5 * rr + 1.5 * alpha - 2.0 * alpha3 - 3.0 * alpha / rr;
          double term1 = C1 * erfc(alpha * r) + C2 * exp(-alpha2 * rr) / rtpi;
          double term2 = C3 * erfc(alpha * r) + C4 * exp(-alpha2 * rr) / rtpi;
          Dtr[0] += term1 + term2 * dx * dx / rr;
          Dtr[1] += term1 + term2 * dy * dy / rr;
          Dtr[2] += term1 + term2 * dy * dy / rr;
        }
      }
}

/* The pair block Dt of calcHI and ewaldBin for the raw difference (dxo, dyo, dzo):
 * every image within rce with the reference's r <= 2 branch, plus the k-sum with
 * cos(k.r) of the raw difference (S/main26.c:1717-1788, 1150-1219). */
static void ref_pair(const orc_lattice_t* lat, double alpha, double rce, int nmax, int kcount, const double* rk,
                     const double* M2, double dxo, double dyo, double dzo, double Dt[9]) {
  const double* L1 = lat->L1;
  const double* L2 = lat->L2;
  const double L = lat->L;
  for (int q = 0; q < 9; ++q) Dt[q] = 0.0;
  for (int nx = -nmax; nx < nmax + 1; nx++)
    for (int ny = -nmax; ny < nmax + 1; ny++)
      for (int nz = -nmax; nz < nmax + 1; nz++) {
        double dx = dxo + nx * L1[0] + ny * L2[0], dy = dyo + nx * L1[1] + ny * L2[1], dz = dzo + nz * L;
        double rr = dx * dx + dy * dy + dz * dz, r = sqrt(rr);
        if (r < rce) {
          double t1, t2;
          orc_rpy_real_coeffs(r, alpha, 1, &t1, &t2);
          double d[3] = {dx, dy, dz};
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Dt[3 * a + b] += (a == b ? t1 : 0.0) + t2 * d[a] * d[b] / rr;
        }
      }
  for (int q = 0; q < kcount; ++q) {
    double coskr = cos(rk[3 * q] * dxo + rk[3 * q + 1] * dyo + rk[3 * q + 2] * dzo);
    for (int a = 0; a < 9; ++a) Dt[a] += M2[9 * q + a] * coskr;
  }
}

/* calcHI restated with its quirks (S/main26.c:1530-1806; X/main14.c:1548-1824). */
void orc_calchi_ref(const orc_lattice_t* lat, int N, const double* rx, const double* ry, const double* rz,
                    uint64_t t, int nt_avg, const orc_bins_t* bins, const double* table, double box_volume,
                    double* M, int* far_count, int* kcount_out) {
  (void)t;
  const int tp = lat->tp, krbc = (lat->flow == 2), treal = lat->treal;
  const double theta = lat->theta, L2yp = lat->L2yp;
  double rce, M_HI = 4.0;
  int nmax;
  if (treal <= tp / 2) {
    rce = 1.0 * L2yp;
    nmax = 1;
  } else {
    rce = 2.0 * L2yp;
    nmax = 2;
  }
  const int nt = (int)round((double)treal / tp * nt_avg) % nt_avg;
  const double alpha = M_HI / rce;
  double *rk, *M2, Dtr[3];
  const int ktotal = ref_kspace(lat, M_HI, rce, box_volume, &rk, &M2);
  if (kcount_out) *kcount_out = ktotal;
  ref_self(lat, alpha, nmax, ktotal, M2, Dtr);
  size_t n3 = (size_t)3 * N;
  memset(M, 0, sizeof(double) * n3 * n3);
  for (int i = 0; i < N; i++) {
    M[(3 * i) * n3 + 3 * i] = Dtr[0];
    M[(3 * i + 1) * n3 + 3 * i + 1] = Dtr[1];
    M[(3 * i + 2) * n3 + 3 * i + 2] = Dtr[2];
  }
  int nfar = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : nfar)
  for (int i = 0; i < N; i++) {
    double rxi = krbc ? rx[i] * cos(theta) - ry[i] * sin(theta) : rx[i];
    double ryi = krbc ? rx[i] * sin(theta) + ry[i] * cos(theta) : ry[i];
    for (int j = i + 1; j < N; j++) {
      double Dt[9] = {0};
      double rxj = krbc ? rx[j] * cos(theta) - ry[j] * sin(theta) : rx[j];
      double ryj = krbc ? rx[j] * sin(theta) + ry[j] * cos(theta) : ry[j];
      double dmin[3];
      orc_nid_bin(lat, rx, ry, rz, j, i, dmin);
      double drrbmin = sqrt(dmin[0] * dmin[0] + dmin[1] * dmin[1] + dmin[2] * dmin[2]);
      if (drrbmin < 12.0 || !table) {
        if (!(drrbmin < 12.0)) nfar++;
        ref_pair(lat, alpha, rce, nmax, ktotal, rk, M2, rxj - rxi, ryj - ryi, rz[j] - rz[i], Dt);
      } else {
        int bin = orc_binij(bins, lat, rx, ry, rz, i, j);
        /* the shear table spans |dx| <= L/2*edot*tp*dt only (S/Initialization.c:269); the
         * reference reads past it for wider pairs.  Clamp so the restatement stays defined. */
        if (bin < 0) bin = 0;
        if (bin >= bins->num_bins) bin = bins->num_bins - 1;
        memcpy(Dt, &table[(size_t)nt * 9 * bins->num_bins + 9 * (size_t)bin], 9 * sizeof(double));
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          M[(3 * i + a) * n3 + 3 * j + b] = Dt[3 * a + b];
          M[(3 * j + b) * n3 + 3 * i + a] = Dt[3 * a + b];
        }
    }
  }
  if (far_count) *far_count = nfar;
  free(rk);
  free(M2);
}

/* ewaldBin restated (S/main26.c:975-1260, X/main14.c:1009-1383): the pair blocks
 * of every pair i < j (chains after i's own first, then within each chain) from
 * the raw position differences, no rotation; incr = Dt[a]/Dtr[a/3], eps += 2 incr
 * and rowsum[3i + a/3], rowsum[3j + a/3] += incr^2 (both updates, as the serial
 * program makes them).  eps is the sum before CATEA's normalisation.  The
 * reference's split is M_HI = 3.5; M_HI is a parameter so that tests can drive
 * the same code to a converged Ewald sum. */
void orc_ewald_bin(const orc_lattice_t* lat, int N, int Nb, const double* rx, const double* ry, const double* rz,
                   double box_volume, double M_HI, double* eps, double* rowsum, double Dtr[3], int* kcount_out) {
  const int tp = lat->tp, treal = lat->treal;
  const double L2yp = lat->L2yp;
  double rce;
  int nmax;
  if (treal <= tp / 2) {
    rce = 1.0 * L2yp;
    nmax = 1;
  } else {
    rce = 2.0 * L2yp;
    nmax = 2;
  }
  const double alpha = M_HI / rce;
  double *rk, *M2;
  const int kc = ref_kspace(lat, M_HI, rce, box_volume, &rk, &M2);
  if (kcount_out) *kcount_out = kc;
  ref_self(lat, alpha, nmax, kc, M2, Dtr);
  double e = 0.0, Dt[9];
  for (int i = 0; i < 3 * N; ++i) rowsum[i] = 0.0;
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < N; i++) {
      const int jb = pass == 0 ? (i / Nb + 1) * Nb : i + 1, je = pass == 0 ? N : (i / Nb + 1) * Nb;
      for (int j = jb; j < je; j++) {
        ref_pair(lat, alpha, rce, nmax, kc, rk, M2, rx[j] - rx[i], ry[j] - ry[i], rz[j] - rz[i], Dt);
        for (int a = 0; a < 9; a++) {
          double incr = Dt[a] / Dtr[a / 3];
          e += 2.0 * incr;
          rowsum[3 * i + a / 3] += incr * incr;
          rowsum[3 * j + a / 3] += incr * incr;
        }
      }
    }
  *eps = e;
  free(rk);
  free(M2);
}

/* gridHI restated (S/gridHI.c:3-244, X/gridHI.c:3-240): for nt = 0..nt_avg-1 the
 * lattice at t = nt*tp/nt_avg, the split M_HI = 6 with rce = L2yp or 2 L2yp, and
 * the pair block at every bin displacement of invbin (the origin bin at the
 * placeholder (0.6, 0.6, 0.6)).  table[nt][9*bin + a]. */
void orc_far_table(const orc_lattice_t* lat0, int nt_avg, const orc_bins_t* b, double box_volume, double* table) {
  orc_lattice_t lat = *lat0;
  const int nbins = b->nbx * b->nby * b->nbz;
  const int origin = b->nhx * b->nby * b->nbz + b->nhy * b->nbz + b->nhz;
  for (int nt = 0; nt < nt_avg; nt++) {
    uint64_t t = (uint64_t)(nt * lat.tp / nt_avg);
    orc_lattice_update(&lat, t);
    int nmax;
    double rce;
    if (lat.treal <= lat.tp / 2) {
      rce = lat.L2yp;
      nmax = 1;
    } else {
      rce = 2.0 * lat.L2yp;
      nmax = 2;
    }
    const double M_HI = 6.0, alpha = M_HI / rce;
    double *rk, *M2;
    const int kc = ref_kspace(&lat, M_HI, rce, box_volume, &rk, &M2);
    for (int i = 0; i < nbins; i++) {
      double dr[3] = {0.6, 0.6, 0.6};
      if (i != origin) {
        int bin = i;
        int bx = (int)(round(bin / (b->nby * b->nbz)));
        bin -= bx * b->nby * b->nbz;
        bx -= b->nhx;
        int by = (int)(round(bin / (b->nbz)));
        bin -= by * b->nbz;
        by -= b->nhy;
        int bz = bin - b->nhz;
        dr[0] = bx * b->bin_size_x;
        dr[1] = by * b->bin_size_y;
        dr[2] = bz * b->bin_size_z;
      }
      ref_pair(&lat, alpha, rce, nmax, kc, rk, M2, dr[0], dr[1], dr[2], table + (size_t)nt * 9 * nbins + 9 * (size_t)i);
    }
    free(rk);
    free(M2);
  }
}

/* CATEA restated (S/main26.c:1317-1381): eps/(9N^2 - 3N), beta from the quadratic
 * root (1 - sqrt(1 - (3N eps^2 - 3N eps)))/(3N eps^2 - 3N eps), and
 * C_i = sqrt(1/(1 + rowsum_i beta^2)) for one sample (the strain-bin running sums
 * are the caller's).  9N^2 - 3N is formed in FP64: the reference's int expression
 * overflows above N = 15,447. */
double orc_catea(int N, double eps_sum, const double* rowsum, double* C) {
  const int ROW = 3 * N;
  double eps = eps_sum / (9.0 * N * N - 3.0 * N);
  double eps2 = eps * eps;
  double btemp = (1.0 - sqrt(1.0 - (ROW * eps2 - ROW * eps))) / (ROW * eps2 - ROW * eps);
  for (int i = 0; i < 3 * N; i++) C[i] = sqrt(1.0 / (1.0 + rowsum[i] * btemp * btemp));
  return btemp;
}

/* ========================================================== Brownian helpers */
void orc_tea_apply(int n, const double* M, double beta, const double* C, const double* R, double* y) {
  size_t n3 = (size_t)3 * n;
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
      for (int j = 0; j < n; ++j)
        for (int b = 0; b < 3; ++b) {
          double m = M[(3 * i + a) * n3 + 3 * j + b];
          double w = (j == i && b == a) ? 1.0 : beta;
          s += m * w * R[3 * j + b];
        }
      y[3 * i + a] = C[3 * i + a] * s;
    }
  }
}

void orc_matvec(int n3, const double* A, const double* x, double* y) {
#pragma omp parallel for if (n3 > 600)
  for (int i = 0; i < n3; ++i) {
    double s = 0.0;
    for (int j = 0; j < n3; ++j) s += A[(size_t)i * n3 + j] * x[j];
    y[i] = s;
  }
}

/* cyclic Jacobi eigen-decomposition of a symmetric matrix (A overwritten, V eigvecs in columns) */
static void jacobi_eig(int n, double* A, double* V, double* w) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[(size_t)i * n + j] = (i == j);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double v = A[(size_t)i * n + j] * A[(size_t)i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[(size_t)p * n + q];
        if (fabs(apq) < 1e-300) continue;
        double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        double tau = (aqq - app) / (2.0 * apq);
        double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = A[(size_t)i * n + i];
}

int orc_sym_sqrt_apply(int n3, const double* A, const double* z, double* y, double* eigmin, double* eigmax) {
  double* B = (double*)malloc(sizeof(double) * (size_t)n3 * n3);
  double* V = (double*)malloc(sizeof(double) * (size_t)n3 * n3);
  double* w = (double*)malloc(sizeof(double) * (size_t)n3);
  double* t = (double*)malloc(sizeof(double) * (size_t)n3);
  memcpy(B, A, sizeof(double) * (size_t)n3 * n3);
  jacobi_eig(n3, B, V, w);
  double lo = 1e300, hi = -1e300;
  for (int k = 0; k < n3; ++k) {
    double s = 0.0;
    for (int i = 0; i < n3; ++i) s += V[(size_t)i * n3 + k] * z[i];
    if (w[k] < lo) lo = w[k];
    if (w[k] > hi) hi = w[k];
    t[k] = s * sqrt(w[k] > 0 ? w[k] : 0.0);
  }
  for (int i = 0; i < n3; ++i) {
    double s = 0.0;
    for (int k = 0; k < n3; ++k) s += V[(size_t)i * n3 + k] * t[k];
    y[i] = s;
  }
  if (eigmin) *eigmin = lo;
  if (eigmax) *eigmax = hi;
  free(B);
  free(V);
  free(w);
  free(t);
  return lo > 0 ? 0 : -1;
}

/* c = beta0 * T^{1/2} e1 for the j x j tridiagonal T (alpha diag, beta off-diag) */
static void tridiag_sqrt_e1(int j, const double* al, const double* be, double beta0, double* c) {
  double* T = (double*)calloc((size_t)j * j, sizeof(double));
  double* V = (double*)malloc(sizeof(double) * (size_t)j * j);
  double* w = (double*)malloc(sizeof(double) * (size_t)j);
  for (int i = 0; i < j; ++i) {
    T[i * j + i] = al[i];
    if (i + 1 < j) T[i * j + i + 1] = T[(i + 1) * j + i] = be[i];
  }
  jacobi_eig(j, T, V, w);
  for (int i = 0; i < j; ++i) {
    double s = 0.0;
    for (int k = 0; k < j; ++k) s += V[i * j + k] * sqrt(w[k] > 0 ? w[k] : 0.0) * V[0 * j + k];
    c[i] = beta0 * s;
  }
  free(T);
  free(V);
  free(w);
}

int orc_lanczos_sqrt(int n3, const double* A, const double* z, double tol, int maxit, double* y, int* iters) {
  double* Vb = (double*)calloc((size_t)(maxit + 1) * n3, sizeof(double));
  double* w = (double*)malloc(sizeof(double) * (size_t)n3);
  double* al = (double*)calloc((size_t)maxit + 1, sizeof(double));
  double* be = (double*)calloc((size_t)maxit + 1, sizeof(double));
  double* c = (double*)calloc((size_t)maxit + 1, sizeof(double));
  double* cp = (double*)calloc((size_t)maxit + 1, sizeof(double));
  double b0 = 0.0;
  for (int i = 0; i < n3; ++i) b0 += z[i] * z[i];
  b0 = sqrt(b0);
  int status = -1, j;
  for (int i = 0; i < n3; ++i) Vb[i] = z[i] / b0;
  for (j = 1; j <= maxit; ++j) {
    double* vj = Vb + (size_t)(j - 1) * n3;
    orc_matvec(n3, A, vj, w);
    double a = 0.0;
    for (int i = 0; i < n3; ++i) a += vj[i] * w[i];
    al[j - 1] = a;
    for (int i = 0; i < n3; ++i) {
      w[i] -= a * vj[i];
      if (j > 1) w[i] -= be[j - 2] * Vb[(size_t)(j - 2) * n3 + i];
    }
    double b = 0.0;
    for (int i = 0; i < n3; ++i) b += w[i] * w[i];
    b = sqrt(b);
    be[j - 1] = b;
    tridiag_sqrt_e1(j, al, be, b0, c);
    if (j > 1) {
      double num = 0.0, den = 0.0;
      for (int i = 0; i < j; ++i) {
        double d = c[i] - (i < j - 1 ? cp[i] : 0.0);
        num += d * d;
        den += c[i] * c[i];
      }
      if (sqrt(num / den) < tol) {
        status = 0;
        break;
      }
    }
    memcpy(cp, c, sizeof(double) * (size_t)j);
    if (j == maxit || b == 0.0) break;
    for (int i = 0; i < n3; ++i) Vb[(size_t)j * n3 + i] = w[i] / b;
  }
  if (j > maxit) j = maxit;
  for (int i = 0; i < n3; ++i) {
    double s = 0.0;
    for (int k = 0; k < j; ++k) s += Vb[(size_t)k * n3 + i] * c[k];
    y[i] = s;
  }
  if (iters) *iters = j;
  free(Vb);
  free(w);
  free(al);
  free(be);
  free(c);
  free(cp);
  return status;
}

/* ========================================================== Charge-hopping MC */
static void rebuild_charge(orc_mc_state_t* s, const int* from) {
  for (int i = 0; i < s->n; ++i) s->charge[i] = 0;
  for (int j = 0; j < s->nq; ++j) s->charge[from[j]] = 1;
}

/* Pair terms of the double loop of chargeHop for unit charges,
 * (lambda_b*1*1/r)*exp(-kappa_d*r), cached per bead pair for one sweep (the
 * positions do not change during a sweep).  A term for charges qi, qj in
 * {-1, 1} is qi*qj times the cached value exactly (round to nearest is
 * symmetric under negation).  Slots are given to beads as they first carry a
 * charge; g holds NaN until a pair is evaluated. */
typedef struct {
  int cap, used;
  int* slot;  /* bead -> slot or -1 */
  double* g;  /* cap x cap */
} pair_cache_t;

static double pair_term(const orc_lattice_t* lat, const double* rx, const double* ry, const double* rz, int i, int j,
                        double lambda_b, double kappa_d) {
  double d[3];
  orc_nid(lat, rx, ry, rz, i, j, d);
  double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double r = sqrt(rr);
  return (lambda_b * 1 * 1 / r) * exp(-kappa_d * r);
}

static int cache_slot(pair_cache_t* c, int bead) {
  if (c->slot[bead] < 0) {
    if (c->used >= c->cap) return -1;
    c->slot[bead] = c->used++;
  }
  return c->slot[bead];
}

/* E = sum over bead pairs i < j, both charged, in the reference's loop order
 * (S/main26.c:1955-1970, 2181-2197).  With a cache the terms of new pairs are
 * evaluated (in parallel) before the sequential sum; without one, in the loop. */
static double mc_energy(orc_mc_state_t* s, const orc_lattice_t* lat, const double* rx, const double* ry,
                        const double* rz, double lambda_b, double kappa_d, pair_cache_t* pc, int* list) {
  double E = 0;
  if (s->nq <= 1) return E;
  int q = 0;
  for (int i = 0; i < s->n; ++i)
    if (s->charge[i] != 0) list[q++] = i;
  int cached = pc != NULL;
  if (cached)
    for (int a = 0; a < q && cached; ++a) cached = cache_slot(pc, list[a]) >= 0;
  if (!cached) {
    for (int a = 0; a < q; ++a)
      for (int b = a + 1; b < q; ++b) {
        int i = list[a], j = list[b];
        E += s->charge[i] * s->charge[j] * pair_term(lat, rx, ry, rz, i, j, lambda_b, kappa_d);
      }
    return E;
  }
  const size_t cap = (size_t)pc->cap;
#pragma omp parallel for schedule(dynamic, 8)
  for (int a = 0; a < q; ++a) {
    double* row = pc->g + (size_t)pc->slot[list[a]] * cap;
    for (int b = a + 1; b < q; ++b) {
      double* gv = row + pc->slot[list[b]];
      if (isnan(*gv)) {
        double t = pair_term(lat, rx, ry, rz, list[a], list[b], lambda_b, kappa_d);
        *gv = t;
        pc->g[(size_t)pc->slot[list[b]] * cap + pc->slot[list[a]]] = t;
      }
    }
  }
  /* the bead list is rewritten in place as slots; charges of unit magnitude */
  int all_pos = 1;
  for (int a = 0; a < q; ++a) {
    all_pos &= s->charge[list[a]] == 1;
    list[a] = pc->slot[list[a]];
  }
  if (all_pos) {
    for (int a = 0; a < q; ++a) {
      const double* row = pc->g + (size_t)list[a] * cap;
      for (int b = a + 1; b < q; ++b) E += row[list[b]];
    }
    return E;
  }
  for (int a = 0; a < q; ++a) list[a] = -1;  /* recover the beads: rescan */
  q = 0;
  for (int i = 0; i < s->n; ++i)
    if (s->charge[i] != 0) list[q++] = i;
  for (int a = 0; a < q; ++a) {
    const double* row = pc->g + (size_t)pc->slot[list[a]] * cap;
    const int ci = s->charge[list[a]];
    for (int b = a + 1; b < q; ++b) E += ci * s->charge[list[b]] * row[pc->slot[list[b]]];
  }
  return E;
}

static void free_cache(pair_cache_t* pc, int* qlist) {
  if (pc) {
    free(pc->slot);
    free(pc->g);
  }
  free(qlist);
}

/* chargeHop (S/main26.c:1926-2390): one MC sweep over the charges, written as
 * the reference's state machine (test/test2 flags, i_values, E_initial fixed for
 * the sweep, lowest-index "pop" target, continue statements). */
int orc_charge_hop(orc_mc_state_t* s, const orc_lattice_t* lat, const double* rx, const double* ry,
                   const double* rz, double lambda_b, double lambda_d, double barrier, double barrier2,
                   const float* u, int nu, int* cursor, int* log, double* e_initial_out) {
  int i, j, k, test, test2, cnumber, cnumber2 = 0, count;
  int NCharges = s->nq, N = s->n, Nb = s->nb;
  double kappa_d = 1 / lambda_d, E_initial, E_final, mm, dice_roll;
#define DRAW() ((*cursor < nu) ? (double)u[(*cursor)++] : (double)(u[0] * 0.0f + 2.0f))
  if (*cursor >= nu) return -2;
  for (j = 0; j < NCharges; ++j)
    for (i = 0; i < NCharges; ++i)
      if (s->track[j] == s->indices[i] && i == j) s->ivalues[i] = j;
  for (i = 0; i < NCharges; ++i) s->indices[i] = s->track[i];
  /* the reference's Charge[] at this point reflects the previous state (track) */
  rebuild_charge(s, s->track);
  /* the pair-term cache: at most nq + NCharges beads carry a charge in a sweep */
  pair_cache_t pcs, *pc = NULL;
  int* qlist = (int*)malloc(sizeof(int) * (size_t)(NCharges + 2));
  if (NCharges <= 5000) {
    pcs.cap = 2 * NCharges + 2;
    pcs.used = 0;
    pcs.slot = (int*)malloc(sizeof(int) * (size_t)N);
    pcs.g = (double*)malloc(sizeof(double) * (size_t)pcs.cap * pcs.cap);
    if (pcs.slot && pcs.g) {
      for (i = 0; i < N; ++i) pcs.slot[i] = -1;
      for (size_t e = 0; e < (size_t)pcs.cap * pcs.cap; ++e) pcs.g[e] = NAN;
      pc = &pcs;
    } else {
      free(pcs.slot);
      free(pcs.g);
    }
  }
  E_initial = mc_energy(s, lat, rx, ry, rz, lambda_b, kappa_d, pc, qlist);
  if (e_initial_out) *e_initial_out = E_initial;
  for (k = 0; k < NCharges; ++k) {
    int draws = 0;
    int* lg = log ? log + 6 * k : NULL;
    if (lg) {
      lg[0] = k;
      lg[1] = s->indices[k];
      lg[2] = -1;
      lg[3] = -1;
      lg[4] = 0;
      lg[5] = 0;
    }
    rebuild_charge(s, s->track);
    if (*cursor >= nu) {
        free_cache(pc, qlist);
        return -2;
      }
    mm = DRAW();
    draws++;
    cnumber = s->indices[k];
    test = 0;
    test2 = 0;
    int skip = 0;
    for (j = 0; j < N; ++j) {
      if (j != cnumber && j != cnumber + 1 && j != cnumber - 1) {
        double d[3];
        orc_nid(lat, rx, ry, rz, cnumber, j, d);
        double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (rr < 4.41) {
          if (mm < .33) {
            cnumber2 = cnumber + 1;
            for (i = 0; i < NCharges; ++i)
              if (cnumber == (N - 1) || cnumber % (Nb) == Nb - 1 || cnumber == Nb - 1 || s->indices[i] == cnumber2) test2 = 1;
            if (test2 != 0) continue;
            for (i = 0; i < NCharges; ++i)
              if (s->indices[i] == cnumber) {
                s->indices[i] = cnumber2;
                test2 = 0;
              }
            test = 1;
          }
          if (mm > 0.66) {
            cnumber2 = cnumber - 1;
            for (i = 0; i < NCharges; ++i)
              if (cnumber2 < 0 || cnumber == 0 || cnumber % Nb == 0 || s->indices[i] == cnumber2) test2 = 1;
            if (test2 != 0) {
              continue;
            } else {
              for (i = 0; i < NCharges; ++i)
                if (s->indices[i] == cnumber) {
                  s->indices[i] = cnumber2;
                  test2 = 0;
                }
              test = 1;
            }
          } else if (mm <= 0.66 && mm >= 0.33) {
            cnumber2 = j;
            for (i = 0; i < NCharges; ++i)
              if (s->indices[i] == cnumber2) test2 = 1;
            if (test2 != 0) {
              continue;
            } else {
              for (i = 0; i < NCharges; ++i)
                if (s->indices[i] == cnumber) {
                  s->indices[i] = cnumber2;
                  test2 = 0;
                }
              test = 2;
            }
          }
          break;
        }
      }
    }
    if (test == 0) {
      if (mm > .5) {
        cnumber2 = cnumber + 1;
        for (i = 0; i < NCharges; ++i)
          if (cnumber == (N - 1) || cnumber % (Nb) == Nb - 1 || cnumber == Nb - 1 || s->indices[i] == cnumber2) test2 = 1;
        if (test2 != 0) {
          skip = 1;
        } else {
          for (i = 0; i < NCharges; ++i)
            if (s->indices[i] == cnumber) {
              s->indices[i] = cnumber2;
              test2 = 0;
            }
        }
      } else {
        cnumber2 = cnumber - 1;
        for (i = 0; i < NCharges; ++i)
          if (cnumber2 < 0 || cnumber == 0 || cnumber % Nb == 0 || s->indices[i] == cnumber2) test2 = 1;
        if (test2 != 0) {
          skip = 1;
        } else {
          for (i = 0; i < NCharges; ++i)
            if (s->indices[i] == cnumber) {
              s->indices[i] = cnumber2;
              test2 = 0;
              test = 0;
            }
        }
      }
    }
    if (skip) {
      if (lg) lg[5] = draws;
      continue;
    }
    if (lg) {
      lg[2] = cnumber2;
      lg[3] = test;
    }
    rebuild_charge(s, s->indices);
    E_final = mc_energy(s, lat, rx, ry, rz, lambda_b, kappa_d, pc, qlist);
    double b = (test == 2) ? barrier2 : barrier;
    int accepted = 0;
    if (1 <= exp((-b - 0.5 * (E_final - E_initial)))) {
      accepted = 1;
    } else {
      if (*cursor >= nu) {
        free_cache(pc, qlist);
        return -2;
      }
      dice_roll = DRAW();
      draws++;
      if (dice_roll <= exp((-b - 0.5 * (E_final - E_initial)))) accepted = 1;
    }
    if (accepted) {
      count = 0;
      for (i = 0; i < NCharges; ++i)
        if (s->track[i] == cnumber && i == s->ivalues[i] && count == 0) {
          count += 1;
          s->old[i] = cnumber;
          s->indices[i] = cnumber2;
          s->track[i] = cnumber2;
        }
    } else {
      count = 0;
      for (i = 0; i < NCharges; ++i)
        if (s->indices[i] == cnumber2 && i == s->ivalues[i] && count == 0) {
          count += 1;
          s->indices[i] = cnumber;
          s->old[i] = cnumber;
        }
    }
    if (lg) {
      lg[4] = accepted;
      lg[5] = draws;
    }
  }
  rebuild_charge(s, s->track);
  free_cache(pc, qlist);
#undef DRAW
  return 0;
}

/* ================================================ reference timestep (timing) */
/* One HI-iteration step of the reference (S/main26.c:65-159 with updateHI
 * S/main26.c:745-857): lattice, PBC, Verlet check, forces, noise, dense M every
 * srpy steps (calcHI with a far-field table lookup), TEA matvec, MC every mstep. */
double orc_ref_run(const orc_ref_run_t* p, double* rx, double* ry, double* rz, int nsteps, int nthreads,
                   int* status, double* t_hi) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double t0 = omp_get_wtime();
#else
  (void)nthreads;
  double t0 = 0;
#endif
  int N = p->nb * p->nc, n3 = 3 * N;
  orc_lattice_t lat;
  orc_lattice_init(&lat, p->flow, p->L, p->edot, p->dt, 1);
  orc_bins_t bins;
  orc_bins_init(&bins, &lat, 2.0);
  double* table = (double*)calloc((size_t)p->nt_avg * 9 * bins.num_bins, sizeof(double));
  double* M = (double*)malloc(sizeof(double) * (size_t)n3 * n3);
  double* fx = (double*)calloc(N, sizeof(double));
  double* fy = (double*)calloc(N, sizeof(double));
  double* fz = (double*)calloc(N, sizeof(double));
  double* dxv = (double*)calloc(N, sizeof(double));
  double* dyv = (double*)calloc(N, sizeof(double));
  double* dzv = (double*)calloc(N, sizeof(double));
  double* R = (double*)calloc(n3, sizeof(double));
  double* du = (double*)calloc(n3, sizeof(double));
  int* nlist = (int*)calloc(N, sizeof(int));
  int stride = N < 512 ? N : 512;
  int* list = (int*)calloc((size_t)N * stride, sizeof(int));
  int* q = (int*)calloc(N, sizeof(int));
  orc_rng_t rng;
  orc_rng_seed(&rng, p->seed);
  orc_mc_state_t mc;
  mc.n = N;
  mc.nb = p->nb;
  mc.nc = p->nc;
  mc.nq = p->nq;
  mc.track = (int*)calloc(p->nq + 1, sizeof(int));
  mc.indices = (int*)calloc(p->nq + 1, sizeof(int));
  mc.old = (int*)calloc(p->nq + 1, sizeof(int));
  mc.ivalues = (int*)calloc(N, sizeof(int));
  mc.charge = (int*)calloc(N, sizeof(int));
  for (int i = 0; i < p->nq; ++i) {
    int c = (int)(orc_ran1(&rng) * N);
    if (q[c] == 0) {
      q[c] = 1;
      mc.track[i] = mc.indices[i] = mc.old[i] = c;
    } else {
      i--;
    }
  }
  double box_volume = p->L * p->L * p->L;
  double pp = sqrt(2.0 / p->dt);
  int st = 0;
  orc_lattice_update(&lat, 0);
  orc_apply_pbc(&lat, N, rx, ry, rz);
  orc_verlet(&lat, N, rx, ry, rz, 6.0, nlist, list, stride);
  int nu = 2 * p->nq + 8;
  float* uu = (float*)malloc(sizeof(float) * nu);
  for (int t = 0; t < nsteps; ++t) {
    for (int i = 0; i < N; ++i) fx[i] = fy[i] = fz[i] = 0.0;
    orc_lattice_update(&lat, (uint64_t)t);
    orc_apply_pbc(&lat, N, rx, ry, rz);
    if (orc_check_verlet(N, dxv, dyv, dzv, 0.5)) {
      orc_verlet(&lat, N, rx, ry, rz, 6.0, nlist, list, stride);
      for (int i = 0; i < N; ++i) dxv[i] = dyv[i] = dzv[i] = 0.0;
    }
    if (orc_bond_force(&lat, p->nb, p->nc, rx, ry, rz, p->spring_fene, p->kappas, p->qmax, 0.0, fx, fy, fz, NULL)) {
      st = -1;
      break;
    }
    orc_lj_force(&lat, N, rx, ry, rz, nlist, list, stride, 25.0, p->epsilon, 0, fx, fy, fz, NULL);
    orc_coulomb_force(&lat, N, rx, ry, rz, q, p->lambda_b, p->lambda_d, fx, fy, fz);
    for (int i = 0; i < n3; ++i) R[i] = (double)orc_gasdev(&rng);
    if (t % p->srpy == 0) {
#ifdef _OPENMP
      double h0 = omp_get_wtime();
#endif
      orc_calchi_ref(&lat, N, rx, ry, rz, (uint64_t)t, p->nt_avg, &bins, table, box_volume, M, NULL, NULL);
#ifdef _OPENMP
      if (t_hi) *t_hi += omp_get_wtime() - h0;
#endif
    }
    double bt = p->beta;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n3; ++i) {
      double s = 0.0;
      const double* row = M + (size_t)i * n3;
      int bi = i / 3;
      for (int j = 0; j < n3; ++j) {
        int bj = j / 3;
        double w = (bi == bj && i == j) ? 1.0 : bt;
        double fj = (j % 3 == 0) ? fx[bj] : (j % 3 == 1) ? fy[bj] : fz[bj];
        s += row[j] * (fj + pp * w * R[j]);
      }
      du[i] = p->dt * s;
    }
    for (int i = 0; i < N; ++i) {
      double ux = (p->flow == 2) ? p->dt * p->edot * rx[i] : p->dt * p->edot * ry[i];
      double uy = (p->flow == 2) ? -p->dt * p->edot * ry[i] : 0.0;
      double dx = du[3 * i] + ux, dy = du[3 * i + 1] + uy, dz = du[3 * i + 2];
      rx[i] += dx;
      ry[i] += dy;
      rz[i] += dz;
      dxv[i] += dx;
      dyv[i] += dy;
      dzv[i] += dz;
    }
    if (p->nq > 0 && p->mstep > 0 && t % p->mstep == 0) {
      int cur = 0;
      orc_ran1_fill(&rng, nu, uu);
      orc_charge_hop(&mc, &lat, rx, ry, rz, p->lambda_b, p->lambda_d, p->barrier, p->barrier2, uu, nu, &cur, NULL, NULL);
      for (int i = 0; i < N; ++i) q[i] = mc.charge[i];
    }
  }
#ifdef _OPENMP
  double el = omp_get_wtime() - t0;
#else
  double el = 0;
#endif
  free(table);
  free(M);
  free(fx);
  free(fy);
  free(fz);
  free(dxv);
  free(dyv);
  free(dzv);
  free(R);
  free(du);
  free(nlist);
  free(list);
  free(q);
  free(mc.track);
  free(mc.indices);
  free(mc.old);
  free(mc.ivalues);
  free(mc.charge);
  free(uu);
  if (status) *status = st;
  return el;
}

/* ============================================================ Metropolis-Hastings hops */
float orc_bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

static double mh_dist(const orc_lattice_t* lat, const double* rx, const double* ry, const double* rz, int i, int j) {
  double d[3];
  orc_nid(lat, rx, ry, rz, i, j, d);
  return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

static double mh_cos_at(const orc_lattice_t* lat, const double* rx, const double* ry, const double* rz, int b,
                        int nb) {
  int pos = b % nb;
  if (pos == 0 || pos == nb - 1) return 0.0;
  double u[3], v[3];
  orc_nid(lat, rx, ry, rz, b - 1, b, u);
  orc_nid(lat, rx, ry, rz, b + 1, b, v);
  double uu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), vv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) / (uu * vv);
}

/* distance between two beads from their vectors relative to a common anchor */
static double mh_rel_dist(const double a[3], const double b[3]) {
  double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

void orc_mh_features(const orc_lattice_t* lat, const orc_mh_params_t* p, const double* rx, const double* ry,
                     const double* rz, const int* qbead, int src, int dst, int qk, int clr, int set, int ncand,
                     int anchor, double f[16]) {
  const int nb = p->nb;
  const double R4 = p->feature_radius, kap = 1.0 / p->lambda_d;
  for (int k = 0; k < 16; ++k) f[k] = 0.0;
  f[0] = mh_dist(lat, rx, ry, rz, src, dst);
  int same = (src / nb) == (dst / nb);
  int gap = abs(dst - src);
  f[1] = (same && gap == 1) ? 1.0 : 0.0;
  f[2] = same ? (double)(gap < 20 ? gap : 20) / 20.0 : 1.0;
  double vs[3], vd[3];
  orc_nid(lat, rx, ry, rz, src, anchor, vs);
  orc_nid(lat, rx, ry, rz, dst, anchor, vd);
  for (int m = 0; m < p->n; ++m) {
    if (m == src || m == dst) continue;
    double vm[3];
    orc_nid(lat, rx, ry, rz, m, anchor, vm);
    double rdm = mh_rel_dist(vd, vm), rsm = mh_rel_dist(vs, vm);
    int qm = (m == clr) ? 0 : (m == set) ? qk : qbead[m];
    if (qm != 0 && (rdm < R4 || rsm < R4))
      f[3] += p->lambda_b * (double)(qk * qm) * (exp(-kap * rdm) / rdm - exp(-kap * rsm) / rsm);
    if (rdm < R4) {
      f[8] += 0.1;
      if (qm != 0) {
        f[4] += 1.0;
        f[6] += (double)(qk * qm);
      }
    }
    if (rsm < R4) {
      f[9] += 0.1;
      if (qm != 0) {
        f[5] += 1.0;
        f[7] += (double)(qk * qm);
      }
    }
  }
  f[10] = mh_cos_at(lat, rx, ry, rz, dst, nb);
  f[11] = mh_cos_at(lat, rx, ry, rz, src, nb);
  f[12] = (dst % nb == 0 || dst % nb == nb - 1) ? 1.0 : 0.0;
  f[13] = (src % nb == 0 || src % nb == nb - 1) ? 1.0 : 0.0;
  f[14] = (double)qk;
  f[15] = (double)ncand / 10.0;
}

double orc_mlp_forward(const float* w, const double f[16]) {
  const float *W1 = w, *b1 = W1 + 64 * 16, *W2 = b1 + 64, *b2 = W2 + 64 * 64, *W3 = b2 + 64, *b3 = W3 + 64;
  double h1[64], h2[64];
  for (int j = 0; j < 64; ++j) {
    double a = b1[j];
    for (int i = 0; i < 16; ++i) a += (double)W1[j * 16 + i] * f[i];
    h1[j] = a > 0 ? a : 0;
  }
  for (int j = 0; j < 64; ++j) {
    double a = b2[j];
    for (int i = 0; i < 64; ++i) a += (double)W2[j * 64 + i] * h1[i];
    h2[j] = a > 0 ? a : 0;
  }
  double s = b3[0];
  for (int j = 0; j < 64; ++j) s += (double)W3[j] * h2[j];
  return s;
}

/* candidates of src: uncharged beads (under clr/set) other than src within hop radius */
static int mh_candidates(const orc_lattice_t* lat, const orc_mh_params_t* p, const double* rx, const double* ry,
                         const double* rz, const int* qbead, int src, int clr, int set, int* out) {
  int nc = 0;
  double R2 = p->hop_radius * p->hop_radius;
  for (int b = 0; b < p->n; ++b) {
    if (b == src) continue;
    int qb = (b == clr) ? 0 : (b == set) ? 1 : qbead[b];
    if (qb != 0) continue;
    double d[3];
    orc_nid(lat, rx, ry, rz, src, b, d);
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < R2) out[nc++] = b;
  }
  return nc;
}

/* proposal probabilities over the candidates (FP64) */
static void mh_proposal(const orc_lattice_t* lat, const orc_mh_params_t* p, const double* rx, const double* ry,
                        const double* rz, const int* qbead, const float* w, int src, int qk, int clr, int set,
                        const int* cand, int nc, int anchor, double* q) {
  if (!w) {
    for (int c = 0; c < nc; ++c) q[c] = 1.0 / nc;
    return;
  }
  double z = 0.0;
  for (int c = 0; c < nc; ++c) {
    double f[16];
    orc_mh_features(lat, p, rx, ry, rz, qbead, src, cand[c], qk, clr, set, nc, anchor, f);
    q[c] = 1.0 / (1.0 + exp(-orc_mlp_forward(w, f)));  /* predicted acceptance: Barker-type weight */
    z += q[c];
  }
  for (int c = 0; c < nc; ++c) q[c] = (1.0 - p->floor_eps) * q[c] / z + p->floor_eps / nc;
}

int orc_mh_sweep(const orc_lattice_t* lat, const orc_mh_params_t* p, const double* rx, const double* ry,
                 const double* rz, int nq, int* track, const int* sign, int* qbead, const float* w,
                 const float* u, int nattempt, int* mlog, float* feats) {
  int* cand = (int*)malloc(sizeof(int) * p->n);
  double* q = (double*)malloc(sizeof(double) * p->n);
  float* wr = NULL;
  int nacc = 0;
  const double kap = 1.0 / p->lambda_d;
  if (w) {
    int np = 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1;
    wr = (float*)malloc(sizeof(float) * np);
    for (int i = 0; i < np; ++i) wr[i] = orc_bf16_round(w[i]);
  }
  for (int a = 0; a < nattempt; ++a) {
    float u0 = u[3 * a], u1 = u[3 * a + 1], u2 = u[3 * a + 2];
    int k = (int)(u0 * (float)nq);
    if (k > nq - 1) k = nq - 1;
    int x = track[k], qk = sign[k];
    int* lg = mlog ? mlog + 6 * a : NULL;
    float* fa = feats ? feats + 16 * a : NULL;
    if (fa)
      for (int i = 0; i < 16; ++i) fa[i] = 0.0f;
    int nc = mh_candidates(lat, p, rx, ry, rz, qbead, x, -1, -1, cand);
    if (nc == 0) {
      if (lg) {
        lg[0] = k; lg[1] = x; lg[2] = -1; lg[3] = -1; lg[4] = 0; lg[5] = 3;
      }
      continue;
    }
    mh_proposal(lat, p, rx, ry, rz, qbead, wr, x, qk, -1, -1, cand, nc, x, q);
    double tot = 0.0;
    for (int c = 0; c < nc; ++c) tot += q[c];
    double t = (double)u1 * tot, cum = 0.0;
    int pick = nc - 1;
    for (int c = 0; c < nc; ++c) {
      cum += q[c];
      if (cum >= t) {
        pick = c;
        break;
      }
    }
    int y = cand[pick];
    double qf = q[pick];
    if (fa) {
      double f[16];
      orc_mh_features(lat, p, rx, ry, rz, qbead, x, y, qk, -1, -1, nc, x, f);
      for (int i = 0; i < 16; ++i) fa[i] = (float)f[i];
    }
    double de = 0.0;
    for (int m = 0; m < nq; ++m) {
      if (m == k) continue;
      int bm = track[m];
      double rym = mh_dist(lat, rx, ry, rz, y, bm), rxm = mh_dist(lat, rx, ry, rz, x, bm);
      de += p->lambda_b * (double)(qk * sign[m]) * (exp(-kap * rym) / rym - exp(-kap * rxm) / rxm);
    }
    /* reverse proposal in the state after the move */
    int nc2 = mh_candidates(lat, p, rx, ry, rz, qbead, y, x, y, cand);
    mh_proposal(lat, p, rx, ry, rz, qbead, wr, y, qk, x, y, cand, nc2, x, q);
    double qr = 0.0;
    for (int c = 0; c < nc2; ++c)
      if (cand[c] == x) qr = q[c];
    double lnrho = -de + log(qr) - log(qf);
    int adj = ((x / p->nb) == (y / p->nb)) && abs(y - x) == 1;
    double ab = exp(-(adj ? p->barrier : p->barrier2));
    double sr = exp(0.5 * lnrho);
    double m1 = ab * sr > ab / sr ? ab * sr : ab / sr;
    double A = ab * sr / (m1 > 1.0 ? m1 : 1.0);
    int acc = (double)u2 < A;
    if (acc) {
      track[k] = y;
      qbead[x] = 0;
      qbead[y] = qk;
      ++nacc;
    }
    if (lg) {
      lg[0] = k; lg[1] = x; lg[2] = y; lg[3] = adj ? 0 : 2; lg[4] = acc; lg[5] = 3;
    }
  }
  free(cand);
  free(q);
  free(wr);
  return nacc;
}
