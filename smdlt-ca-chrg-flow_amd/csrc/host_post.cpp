// Host-side pieces of the reference's run around the timestep: the TEA
// strain-bin averages of iteration 0 and their hand-off to iteration 1 (CATEA's
// bookkeeping and decomp/B file, printTEA's decomp/DC file, resetAverage, the
// DC reader), and the mean-squared displacement post-processing (calcMSD,
// calcMSDcom).  The reference runs these on the CPU as well; they keep its
// file formats, operation order and index layout so that its analysis reads
// the files unchanged.
#pragma STDC FP_CONTRACT OFF
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "smdlt_hip.h"

struct smdlt_tea_avg {
  int nb, nc, sbins;
  double strain_interval;
  std::vector<double> brun;  // [sbins + 1]
  std::vector<double> crun;  // [sbins + 1][3 nb]
  std::vector<int> count;    // [sbins + 1]
};

namespace {
int strain_bin(double strain, double interval, int sbins) {
  int b = (int)floor(strain / interval);
  return b > sbins ? sbins : b;  // steady-state bin (S/main26.c:1346-1349)
}

// calcMSD / calcMSDcom's windowed sums (S/main26.c:2584-2629, X/main14.c:3172-3212)
// over the sorted displacement layout: for each shift i and series q < nq, the
// windows j = j0 .. wmax - i - 1 of i consecutive entries starting at
// j + q * Tmax (the reference's index, which does not follow the sort's
// stride Tmax + 1 for q > 0).
void msd_windows(const std::vector<double>& vx, const std::vector<double>& vy, const std::vector<double>& vz, int nq,
                 int Tmax, int j0, int ishift_end, double* out4) {
  const int wmax = Tmax + 1;
  for (int i = 1; i < ishift_end; ++i) {
    double avg = 0, avgx = 0, avgy = 0, avgz = 0;
    for (int q = 0; q < nq; ++q) {
      double R = 0, Rx = 0, Ry = 0, Rz = 0, cnt = 0;
      for (int j = j0; j < wmax - i; ++j) {
        double dx = 0.0, dy = 0.0, dz = 0.0;
        for (int k = 0; k < i; ++k) {
          dx += vx[j + k + q * Tmax];
          dy += vy[j + k + q * Tmax];
          dz += vz[j + k + q * Tmax];
        }
        R += dx * dx + dy * dy + dz * dz;
        Rx += dx * dx;
        Ry += dy * dy;
        Rz += dz * dz;
        cnt++;
      }
      avg += R / cnt;
      avgx += Rx / cnt;
      avgy += Ry / cnt;
      avgz += Rz / cnt;
    }
    double* o = out4 + 4 * (i - 1);
    o[0] = avg;
    o[1] = avgx;
    o[2] = avgy;
    o[3] = avgz;
  }
}

// The sort of calcMSD (S/main26.c:2548-2561): series j's entries 1..Tmax go to
// j * Tmax + i for j = 0 and to j * (Tmax + 1) + i otherwise; the entries the
// sort does not reach stay zero.
void msd_sort(const std::vector<double>& w, int nser, int Tmax, std::vector<double>& v) {
  v.assign((size_t)nser * (Tmax + 1) + Tmax + 1, 0.0);
  for (int j = 0; j < nser; ++j)
    for (int i = 1; i < Tmax + 1; ++i) {
      const int ind1 = (j > 0) ? j * (Tmax + 1) + i : j * Tmax + i;
      v[ind1] = w[(size_t)i * nser + j];
    }
}
}  // namespace

extern "C" {

int smdlt_tea_avg_create(int nb, int nc, int sbins, double strain_interval, smdlt_tea_avg** out) {
  if (!out || nb <= 0 || nc <= 0 || sbins < 0 || !(strain_interval > 0)) return SMDLT_E_ARG;
  smdlt_tea_avg* a = new smdlt_tea_avg();
  a->nb = nb;
  a->nc = nc;
  a->sbins = sbins;
  a->strain_interval = strain_interval;
  a->brun.assign(sbins + 1, 0.0);
  a->crun.assign((size_t)(sbins + 1) * 3 * nb, 0.0);
  a->count.assign(sbins + 1, 0);
  *out = a;
  return 0;
}

void smdlt_tea_avg_destroy(smdlt_tea_avg* a) { delete a; }

int smdlt_tea_avg_add(smdlt_tea_avg* a, uint64_t t, double strain, double beta, const double* C,
                      const char* beta_path) {
  if (!a || !C || !(strain >= 0)) return SMDLT_E_ARG;
  const int n = a->nb * a->nc, row = 3 * a->nb;
  const int b = strain_bin(strain, a->strain_interval, a->sbins);
  a->brun[b] += beta;
  double* cr = a->crun.data() + (size_t)b * row;
  for (int i = 0; i < 3 * n; ++i) cr[i % row] += C[i];
  a->count[b]++;
  if (beta_path) {  // CATEA's decomp/B record (S/main26.c:1371-1376)
    FILE* f = fopen(beta_path, "a");
    if (!f) return SMDLT_E_ARG;
    fprintf(f, "%lu %lf\n", (unsigned long)t, beta);
    for (int i = 0; i < n; ++i) fprintf(f, "%d %lf %lf %lf\n", i, C[3 * i], C[3 * i + 1], C[3 * i + 2]);
    if (fclose(f) != 0) return SMDLT_E_ARG;
  }
  return b;
}

int smdlt_tea_avg_write(const smdlt_tea_avg* a, uint64_t t, const char* path) {
  if (!a || !path) return SMDLT_E_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return SMDLT_E_ARG;
  fprintf(f, "%lu\n", (unsigned long)t);
  const int row = 3 * a->nb;
  for (int i = 0; i < a->sbins + 1; ++i) {
    const int n = a->count[i];
    if (n > 0) {
      fprintf(f, "%d %d %.12e\n", i, n, a->brun[i] / n);
      const double* cr = a->crun.data() + (size_t)i * row;
      for (int j = 0; j < a->nb; j++)
        fprintf(f, "%d %.12e %.12e %.12e\n", j, cr[3 * j] / (a->nc * n), cr[3 * j + 1] / (a->nc * n),
                cr[3 * j + 2] / (a->nc * n));
    } else {
      fprintf(f, "%d %d %.12e\n", i, 0, 0.0);
      for (int j = 0; j < a->nb; ++j) fprintf(f, "%d %.12e %.12e %.12e\n", j, 0.0, 0.0, 0.0);
    }
  }
  return fclose(f) == 0 ? 0 : SMDLT_E_ARG;
}

int smdlt_tea_avg_finish(smdlt_tea_avg* a, double* beta, double* C, int* counts) {
  if (!a) return SMDLT_E_ARG;
  const int row = 3 * a->nb;
  for (int i = 0; i < a->sbins + 1; ++i) {
    const int n = a->count[i];
    if (beta) beta[i] = a->brun[i] / n;  // an empty bin gives 0/0, as the reference's resetAverage
    if (C)
      for (int j = 0; j < row; j++) C[(size_t)i * row + j] = a->crun[(size_t)i * row + j] / (a->nc * n);
    if (counts) counts[i] = n;
    a->brun[i] = 0.0;
    for (int j = 0; j < row; j++) a->crun[(size_t)i * row + j] = 0.0;
    a->count[i] = 0;
  }
  return 0;
}

int smdlt_tea_read(const char* path, int nb, int sbins, double* beta, double* C, uint64_t* t) {
  if (!path || nb <= 0 || sbins < 0 || !beta || !C) return SMDLT_E_ARG;
  FILE* f = fopen(path, "r");
  if (!f) return SMDLT_E_ARG;
  unsigned long tt = 0;
  int ok = fscanf(f, "%lu\n", &tt) == 1;
  for (int i = 0; ok && i < sbins + 1; ++i) {
    int a = 0, b = 0;
    ok = fscanf(f, "%d %d %le\n", &a, &b, &beta[i]) == 3;
    for (int j = 0; ok && j < nb; j++)
      ok = fscanf(f, "%d %le %le %le\n", &a, &C[(size_t)i * 3 * nb + 3 * j], &C[(size_t)i * 3 * nb + 3 * j + 1],
                  &C[(size_t)i * 3 * nb + 3 * j + 2]) == 4;
  }
  fclose(f);
  if (!ok) return SMDLT_E_ARG;
  if (t) *t = tt;
  return 0;
}

int smdlt_tea_coefficients(int nb, int nc, int sbins, double strain_interval, double strain, const double* beta_bins,
                           const double* C_bins, double* beta, double* C) {
  if (nb <= 0 || nc <= 0 || sbins < 0 || !(strain_interval > 0) || !(strain >= 0) || !beta_bins || !C_bins || !beta ||
      !C)
    return SMDLT_E_ARG;
  const int b = strain_bin(strain, strain_interval, sbins), row = 3 * nb;
  *beta = beta_bins[b];
  for (int i = 0; i < 3 * nb * nc; ++i) C[i] = C_bins[(size_t)b * row + i % row];
  return b;
}

int smdlt_msd(const char* ch_path, int ncharges, int tsteps, int msd_start, double dt, int printprops,
              const char* out_path, double* out, int cap) {
  if (!ch_path || ncharges <= 0 || tsteps <= 0 || msd_start < 0 || printprops <= 0) return SMDLT_E_ARG;
  const int Tmax = tsteps, wmax = Tmax + 1, nsamp = ncharges * wmax;
  std::vector<double> wx(nsamp), wy(nsamp), wz(nsamp);
  FILE* f = fopen(ch_path, "r");
  if (!f) return SMDLT_E_ARG;
  for (int i = 0; i < nsamp; i++) {
    int a, b, c, d;
    if (fscanf(f, "%d %d %d %d %lf %lf %lf\n", &a, &b, &c, &d, &wx[i], &wy[i], &wz[i]) != 7) {
      fclose(f);
      return SMDLT_E_ARG;
    }
  }
  fclose(f);
  std::vector<double> vx, vy, vz;
  msd_sort(wx, ncharges, Tmax, vx);
  msd_sort(wy, ncharges, Tmax, vy);
  msd_sort(wz, ncharges, Tmax, vz);
  const int nq = ncharges >= 25 ? 25 : ncharges;
  const int rows = wmax - msd_start - 1;
  if (rows <= 0) return 0;
  if ((size_t)(msd_start + nq * Tmax) > vx.size()) return SMDLT_E_ARG;
  std::vector<double> r4(4 * (size_t)rows);
  msd_windows(vx, vy, vz, nq, Tmax, msd_start, wmax - msd_start, r4.data());
  FILE* o = out_path ? fopen(out_path, "w") : nullptr;
  if (out_path && !o) return SMDLT_E_ARG;
  for (int i = 1; i < wmax - msd_start; i++) {
    const double* v = r4.data() + 4 * (i - 1);
    const double tau = i * dt * printprops;
    if (o) fprintf(o, "%f,%lf,%lf,%lf,%lf\n", tau, v[0] / nq, v[1] / nq, v[2] / nq, v[3] / nq);
    if (out && i - 1 < cap) {
      double* w = out + 5 * (i - 1);
      w[0] = tau;
      for (int k = 0; k < 4; ++k) w[1 + k] = v[k] / nq;
    }
  }
  if (o && fclose(o) != 0) return SMDLT_E_ARG;
  return rows;
}

int smdlt_msd_com(const char* com_path, int tsteps, double dt, int printprops, const char* out_path, double* out,
                  int cap) {
  if (!com_path || tsteps <= 1 || printprops <= 0) return SMDLT_E_ARG;
  const int Tmax = tsteps, wmax = Tmax + 1, nsamp = wmax;
  std::vector<double> wx(nsamp), wy(nsamp), wz(nsamp);
  FILE* f = fopen(com_path, "r");
  if (!f) return SMDLT_E_ARG;
  for (int i = 0; i < nsamp; i++) {
    int a;
    if (fscanf(f, "%d %lf %lf %lf\n", &a, &wx[i], &wy[i], &wz[i]) != 4) {
      fclose(f);
      return SMDLT_E_ARG;
    }
  }
  fclose(f);
  std::vector<double> vx, vy, vz;
  msd_sort(wx, 1, Tmax, vx);
  msd_sort(wy, 1, Tmax, vy);
  msd_sort(wz, 1, Tmax, vz);
  const int rows = wmax - 2;
  std::vector<double> r4(4 * (size_t)rows);
  msd_windows(vx, vy, vz, 1, Tmax, 1, wmax - 1, r4.data());
  FILE* o = out_path ? fopen(out_path, "w") : nullptr;
  if (out_path && !o) return SMDLT_E_ARG;
  for (int i = 1; i < wmax - 1; i++) {
    const double* v = r4.data() + 4 * (i - 1);
    const double tau = i * dt * printprops;
    if (o) fprintf(o, "%f,%lf,%lf,%lf,%lf\n", tau, v[0], v[1], v[2], v[3]);
    if (out && i - 1 < cap) {
      double* w = out + 5 * (i - 1);
      w[0] = tau;
      for (int k = 0; k < 4; ++k) w[1 + k] = v[k];
    }
  }
  if (o && fclose(o) != 0) return SMDLT_E_ARG;
  return rows;
}

}  // extern "C"
