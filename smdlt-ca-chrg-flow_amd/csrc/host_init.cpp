// Host-side set-up of a run as the reference does it: its random stream
// (ran1, gasdev), the initial chains and charges of initChains, and the restart
// reader.  These run on the CPU in the reference too; they are here so that a
// driver on this library reproduces the reference's initial state and noise
// stream from the same seed.  No contraction: every value follows the
// reference's operation order in IEEE double/float.
#pragma STDC FP_CONTRACT OFF
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "smdlt_hip.h"

namespace {
// ran1's constants (S/Parameters.h:21-34): one Schrage LCG (a = 40014,
// m = 2147483563) feeding a 32-entry Bays-Durham shuffle.
constexpr int64_t kM1 = 2147483563, kA1 = 40014, kQ1 = 53668, kR1 = 12211;
constexpr int64_t kM2 = 2147483399, kQ2 = 52774;
constexpr int kTab = 32;
constexpr int64_t kDiv = 1 + (kM1 - 1) / kTab;
constexpr double kScale = 1.0 / kM1, kMax = 1.0 - 1.2e-7;

inline int64_t lcg_next(int64_t x) {
  const int64_t k = x / kQ1;
  x = kA1 * (x - k * kQ1) - k * kR1;
  return x < 0 ? x + kM1 : x;
}
}  // namespace

extern "C" {

void smdlt_ran1_seed(smdlt_ran1* s, int64_t seed) {
  std::memset(s, 0, sizeof(*s));
  s->idum = seed;
  s->idum2 = 123456789;  // the static initial values of ran1 and gasdev
}

// ran1 (S/Initialization.c:601-634).  A non-positive idum refills the shuffle
// table (a positive seed starts from the all-zero table, as the reference's
// statics do), and the second generator stays at its start value: the
// reference computes its quotient but never advances it.
float smdlt_ran1_next(smdlt_ran1* s) {
  if (s->idum <= 0) {
    s->idum = (-(s->idum) < 1) ? 1 : -(s->idum);
    s->idum2 = s->idum;
    for (int j = kTab + 7; j >= 0; --j) {
      s->idum = lcg_next(s->idum);
      if (j < kTab) s->iv[j] = s->idum;
    }
    s->iy = s->iv[0];
  }
  s->idum = lcg_next(s->idum);
  if (s->idum < 0) s->idum2 += kM2;  // never taken (idum >= 0 here), kept for the record
  (void)kQ2;
  const int j = (int)(s->iy / kDiv);
  s->iy = s->iv[j] - s->idum2;
  s->iv[j] = s->idum;
  if (s->iy < 1) s->iy += kM1 - 1;
  const float u = (float)(kScale * (double)s->iy);
  return ((double)u > kMax) ? (float)kMax : u;
}

void smdlt_ran1_fill(smdlt_ran1* s, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = smdlt_ran1_next(s);
}

// gasdev (S/main26.c:515-540): polar Box-Muller in float (log and sqrt in
// double), the second value of each pair kept for the next call; getNoise
// (S/main26.c:708-713) stores the floats in a double array.
void smdlt_gasdev_fill(smdlt_ran1* s, int n, double* out) {
  for (int i = 0; i < n; ++i) {
    if (s->idum < 0) s->iset = 0;
    if (s->iset == 0) {
      float v1, v2, rsq;
      do {
        v1 = (float)(2.0 * (double)smdlt_ran1_next(s) - 1.0);
        v2 = (float)(2.0 * (double)smdlt_ran1_next(s) - 1.0);
        rsq = v1 * v1 + v2 * v2;
      } while (rsq >= 1.0 || rsq == 0.0);
      const float fac = (float)std::sqrt(-2.0 * std::log((double)rsq) / (double)rsq);
      s->gset = v1 * fac;
      s->iset = 1;
      out[i] = (double)(v2 * fac);
    } else {
      s->iset = 0;
      out[i] = (double)s->gset;
    }
  }
}

// initChains for a new run (S/main26.c:346-470): chains placed one after the
// other in the cube [-L/2, L/2)^3.  A chain's first bead is drawn uniformly
// (at the origin when there is one chain) and kept when it is at least sqrt(10)
// from every earlier bead; each further bead sits 2.05 from its predecessor in
// a uniform direction (theta, then phi = acos(2u - 1)), wrapped into the cube,
// and is kept when it is at least 2 from every earlier bead (cube minimum
// image).  After more than 1000 attempts within one chain the whole
// construction restarts from the first chain, the stream continuing.
int smdlt_init_chains(int nb, int nc, double L, smdlt_ran1* s, double* rx, double* ry, double* rz) {
  if (nb < 1 || nc < 1 || !(L > 0.0) || !s || !rx || !ry || !rz) return SMDLT_E_ARG;
  const double half = L / 2;
  auto far_enough = [&](int idx, double min2) {
    for (int k = 0; k < idx; ++k) {
      double dx = rx[idx] - rx[k], dy = ry[idx] - ry[k], dz = rz[idx] - rz[k];
      dx -= L * std::round(dx / L);
      dy -= L * std::round(dy / L);
      dz -= L * std::round(dz / L);
      if (dx * dx + dy * dy + dz * dz < min2) return false;
    }
    return true;
  };
  for (long restarts = 0; restarts < 100000; ++restarts) {
    bool ok = true;
    for (int i = 0; i < nc && ok; ++i) {
      int tries = 0;
      const int index = nb * i;
      for (;;) {
        if (nc == 1) {
          rx[index] = ry[index] = rz[index] = 0.0;
        } else {
          rx[index] = smdlt_ran1_next(s) * L - half;
          ry[index] = smdlt_ran1_next(s) * L - half;
          rz[index] = smdlt_ran1_next(s) * L - half;
        }
        const bool good = far_enough(index, 10.0);
        if (++tries > 1000) {  // counted before the test is looked at, as the reference does
          ok = false;
          break;
        }
        if (good) break;
      }
      for (int j = 1; j < nb && ok; ++j) {
        const int m = index + j;
        for (;;) {
          const double theta = smdlt_ran1_next(s) * 2.0 * M_PI;
          const double phi = std::acos(2.0 * smdlt_ran1_next(s) - 1.0);
          rx[m] = rx[m - 1] + 2.05 * std::sin(phi) * std::cos(theta);
          ry[m] = ry[m - 1] + 2.05 * std::sin(phi) * std::sin(theta);
          rz[m] = rz[m - 1] + 2.05 * std::cos(phi);
          rx[m] -= L * std::round(rx[m] / L);
          ry[m] -= L * std::round(ry[m] / L);
          rz[m] -= L * std::round(rz[m] / L);
          const bool good = far_enough(m, 4.0);
          if (++tries > 1000) {
            ok = false;
            break;
          }
          if (good) break;
        }
      }
    }
    if (ok) return 0;
  }
  return SMDLT_E_GEOMETRY;  // the box is too full for the construction
}

// The charges of initChains (S/main26.c:485-500): ncharges distinct beads, each
// drawn as (int)(ran1 * N) in float and redrawn when already charged.  For N
// above 2^22 the float product of ran1's largest value and N can round to N;
// such a draw is also redrawn (the reference would index past its arrays).
int smdlt_init_charges(int n, int ncharges, smdlt_ran1* s, int* track) {
  if (n < 1 || ncharges < 0 || ncharges > n || !s || (!track && ncharges)) return SMDLT_E_ARG;
  std::vector<char> charged(n, 0);
  for (int i = 0; i < ncharges; ++i) {
    int c;
    do {
      c = (int)(smdlt_ran1_next(s) * (float)n);
    } while (c >= n || charged[c]);
    charged[c] = 1;
    track[i] = c;
  }
  return 0;
}

// The restart reader of initChains (S/main26.c:264-275): frames of "N", "t"
// and N lines "A x y z"; the last complete frame is kept.  Returns the bead
// count of that frame (an error when it differs from n).
int smdlt_read_restart(const char* path, int n, double* rx, double* ry, double* rz, uint64_t* t) {
  if (!path || n < 1 || !rx || !ry || !rz) return SMDLT_E_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return SMDLT_E_ARG;
  std::vector<double> buf(3 * (size_t)n);
  int got = 0, frames = 0;
  unsigned long long tt = 0;
  for (;;) {
    int m;
    unsigned long long ts;
    if (std::fscanf(f, "%d %llu", &m, &ts) != 2) break;
    if (m != n) {
      std::fclose(f);
      return SMDLT_E_ARG;
    }
    int j = 0;
    for (; j < m; ++j)
      if (std::fscanf(f, " A %lf %lf %lf", &buf[3 * j], &buf[3 * j + 1], &buf[3 * j + 2]) != 3) break;
    if (j < m) break;
    for (int k = 0; k < n; ++k) {
      rx[k] = buf[3 * k];
      ry[k] = buf[3 * k + 1];
      rz[k] = buf[3 * k + 2];
    }
    tt = ts;
    got = m;
    ++frames;
  }
  std::fclose(f);
  if (!frames) return SMDLT_E_ARG;
  if (t) *t = (uint64_t)tt;
  return got;
}

}  // extern "C"
