// Charge-hopping Monte Carlo (chargeHop, S/main26.c:1926-2390; X/main14.c:1945-2409)
// on the device: one workgroup per replica walks the charges in order (the
// moves are sequential by definition) while the lanes share the O(N) scan for
// the first nearby non-adjacent bead and the O(Nq) screened-Coulomb energy
// change of each proposal.
//
// Reference mode keeps the reference's proposal state machine and acceptance
// exp(-b - (E_final - E_initial)/2) with E_initial frozen for the sweep.  With
// Charge_track == Charge_indices at every trial boundary (which holds for the
// reference because i_values[i] = i after initChains sets both arrays equal,
// S/main26.c:490-506), the state machine reduces to:
//   mm = U;  j* = lowest j not in {c-1,c,c+1} with |getNID(c,j)|^2 < 4.41
//   j* found:  mm<.33 -> right, mm>.66 -> left, else pop to j* (barrier2)
//   none:      mm>.5  -> right, else left
//   blocked (chain end or occupied target) -> skip without a second draw
//   x = exp(-b - (E(proposed) - E_initial)/2); accept if 1 <= x, else draw U <= x.
// E(proposed) is kept incrementally (E_cur + dE_local) in FP64.
#include "common.h"

#define MC_THREADS 256

__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = warp_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  __syncthreads();
  return s;
}

__device__ __forceinline__ int block_min_int(int v, int* sh) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  int s = INT_MAX;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s = min(s, sh[k]);
  __syncthreads();
  return s;
}

__device__ __forceinline__ double pair_energy(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                              int a, int b, int qa, int qb) {
  double d[3];
  ref_nid(R, rx[a], ry[a], rz[a], rx[b], ry[b], rz[b], d[0], d[1], d[2]);
  double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double r = sqrt(rr);
  return (R.lambda_b * (double)(qa * qb) / r) * exp(-R.kappa_d * r);
}

// energy of all charge pairs of the replica (the double loop of chargeHop,
// S/main26.c:1955-1970, over charged beads only)
__device__ double charge_energy(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                const int* track, const int* sign, double* sh) {
  double e = 0.0;
  const int nq = R.nq;
  for (int a = 0; a < nq - 1; ++a) {
    int as = R.slot0 + track[R.q0 + a], qa = sign[R.q0 + a];
    for (int b = a + 1 + threadIdx.x; b < nq; b += blockDim.x)
      e += pair_energy(R, rx, ry, rz, as, R.slot0 + track[R.q0 + b], qa, sign[R.q0 + b]);
  }
  return block_sum(e, sh);
}

__device__ __forceinline__ float dev_uniform(uint64_t seed, uint64_t sweep, uint32_t draw) {
  uint32_t c[4] = {(uint32_t)sweep, (uint32_t)(sweep >> 32), draw, 0x4D43u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return ((float)(c[0] >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

struct McShared {
  double dsum[MC_THREADS / 64];
  int imin[MC_THREADS / 64];
  int target, kind, skip, accept;
  double x;
};

__global__ __launch_bounds__(MC_THREADS) void k_mc_sweep(const RepDev* __restrict__ reps, int rep_begin,
                                                         const double* __restrict__ rx, const double* __restrict__ ry,
                                                         const double* __restrict__ rz, int* track, int* old,
                                                         const int* sign, unsigned char* occ, long long occ_stride,
                                                         const float* uniforms, int nu, const uint64_t* seeds,
                                                         uint64_t sweep_id, int* consumed, int* log,
                                                         const int* log_off, int* err) {
  int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  __shared__ McShared S;
  const int N = R.n, Nb = R.nb, nq = R.nq;
  unsigned char* oc = occ + (size_t)blockIdx.x * occ_stride;
  for (int i = threadIdx.x; i < N; i += blockDim.x) oc[i] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < nq; k += blockDim.x) oc[track[R.q0 + k]] = 1;
  __syncthreads();
  const double E_initial = charge_energy(R, rx, ry, rz, track, sign, S.dsum);
  double E_cur = E_initial;
  int cursor = 0;
  uint32_t draws = 0;
  const float* ur = uniforms ? uniforms + (size_t)r * nu : nullptr;
  for (int k = 0; k < nq; ++k) {
    const int c = track[R.q0 + k];
    // draw mm
    float mmf;
    if (ur) {
      if (cursor >= nu) {
        if (threadIdx.x == 0) atomicOr(err, 2);
        break;
      }
      mmf = ur[cursor++];
    } else {
      mmf = dev_uniform(seeds[r], sweep_id, draws++);
    }
    const double mm = (double)mmf;
    // first j (ascending) not in {c-1,c,c+1} with |getNID(c,j)|^2 < 4.41
    int jloc = INT_MAX;
    {
      const int cs = R.slot0 + c;
      const double xc = rx[cs], yc = ry[cs], zc = rz[cs];
      for (int j = threadIdx.x; j < N && jloc == INT_MAX; j += blockDim.x) {
        if (j == c || j == c + 1 || j == c - 1) continue;
        double d[3];
        int js = R.slot0 + j;
        ref_nid(R, xc, yc, zc, rx[js], ry[js], rz[js], d[0], d[1], d[2]);
        double rr;
        {
#pragma clang fp contract(off)
          rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        }
        if (rr < 4.41) jloc = j;
      }
    }
    int jfirst = block_min_int(jloc, S.imin);
    if (threadIdx.x == 0) {
      int target, kind;
      bool right_block = (c == N - 1) || (c % Nb == Nb - 1) || (c == Nb - 1) || (c + 1 < N && oc[c + 1]);
      bool left_block = (c - 1 < 0) || (c == 0) || (c % Nb == 0) || (c - 1 >= 0 && oc[c - 1]);
      bool blocked;
      if (jfirst != INT_MAX) {
        if (mm < .33) {
          target = c + 1;
          kind = 1;
          blocked = right_block;
        } else if (mm > 0.66) {
          target = c - 1;
          kind = 1;
          blocked = left_block;
        } else {
          target = jfirst;
          kind = 2;
          blocked = oc[jfirst] != 0;
        }
      } else {
        kind = 0;
        if (mm > .5) {
          target = c + 1;
          blocked = right_block;
        } else {
          target = c - 1;
          blocked = left_block;
        }
      }
      S.target = target;
      S.kind = kind;
      S.skip = blocked ? 1 : 0;
    }
    __syncthreads();
    const int target = S.target, kind = S.kind;
    int* lg = log ? log + 6 * (size_t)(log_off[r] + k) : nullptr;
    if (S.skip) {
      if (threadIdx.x == 0 && lg) {
        lg[0] = k;
        lg[1] = c;
        lg[2] = -1;
        lg[3] = -1;
        lg[4] = 0;
        lg[5] = 1;
      }
      __syncthreads();
      continue;
    }
    // energy change of moving charge k from c to target
    double de = 0.0;
    {
      const int qk = sign[R.q0 + k];
      const int ts = R.slot0 + target, cs = R.slot0 + c;
      for (int b = threadIdx.x; b < nq; b += blockDim.x) {
        if (b == k) continue;
        int bs = R.slot0 + track[R.q0 + b];
        int qb = sign[R.q0 + b];
        de += pair_energy(R, rx, ry, rz, ts, bs, qk, qb) - pair_energy(R, rx, ry, rz, cs, bs, qk, qb);
      }
    }
    de = block_sum(de, S.dsum);
    if (threadIdx.x == 0) {
      double E_final = (nq > 1) ? E_cur + de : 0.0;
      double b = (kind == 2) ? R.barrier2 : R.barrier;
      double x = exp((-b - 0.5 * (E_final - (nq > 1 ? E_initial : 0.0))));
      int ndraw = 1, acc = 0;
      if (1 <= x) {
        acc = 1;
      } else {
        float dice;
        bool ok = true;
        if (ur) {
          if (cursor >= nu) {
            atomicOr(err, 2);
            ok = false;
            dice = 2.0f;
          } else {
            dice = ur[cursor++];
          }
        } else {
          dice = dev_uniform(seeds[r], sweep_id, draws++);
        }
        ndraw = 2;
        if (ok && (double)dice <= x) acc = 1;
      }
      S.accept = acc;
      if (acc) {
        track[R.q0 + k] = target;
        old[R.q0 + k] = c;
        oc[c] = 0;
        oc[target] = 1;
      } else {
        old[R.q0 + k] = c;
      }
      if (lg) {
        lg[0] = k;
        lg[1] = c;
        lg[2] = target;
        lg[3] = kind;
        lg[4] = acc;
        lg[5] = ndraw;
      }
    }
    __syncthreads();
    if (S.accept) E_cur += de;
    // (consumption of the dice draw is mirrored by thread 0's cursor)
    if (ur) {
      // broadcast cursor from thread 0
      if (threadIdx.x == 0) S.kind = cursor;
      __syncthreads();
      cursor = S.kind;
    } else {
      if (threadIdx.x == 0) S.kind = (int)draws;
      __syncthreads();
      draws = (uint32_t)S.kind;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && consumed) consumed[r] = ur ? cursor : (int)draws;
}

void launch_mc_sweep(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* rx, const double* ry,
                     const double* rz, int* track, int* old, const int* sign, unsigned char* occ,
                     long long occ_stride, const float* uniforms, int nu, const uint64_t* seeds, uint64_t sweep_id,
                     int* consumed, int* log, const int* log_off, int* err) {
  if (nrep <= 0) return;
  k_mc_sweep<<<nrep, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, track, old, sign, occ, occ_stride, uniforms, nu,
                                          seeds, sweep_id, consumed, log, log_off, err);
}

__global__ __launch_bounds__(MC_THREADS) void k_mc_energy(const RepDev* __restrict__ reps, int rep_begin,
                                                          const double* rx, const double* ry, const double* rz,
                                                          const int* track, const int* sign, double* e) {
  int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  __shared__ double sh[MC_THREADS / 64];
  double v = (R.nq > 1) ? charge_energy(R, rx, ry, rz, track, sign, sh) : 0.0;
  if (threadIdx.x == 0) e[r] = v;
}

void launch_mc_energy(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* rx, const double* ry,
                      const double* rz, const int* track, const int* sign, double* e) {
  if (nrep > 0) k_mc_energy<<<nrep, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, track, sign, e);
}
