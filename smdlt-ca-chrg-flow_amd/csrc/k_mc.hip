// Charge-hopping Monte Carlo (chargeHop, S/main26.c:1926-2390; X/main14.c:1945-2409)
// on the device.  The moves are sequential by definition; the sweep computes
// what does not depend on earlier moves for all charges at once and decides
// the charges in blocked rounds (see k_mc_prep / k_mc_round / k_mc_fold).
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
// E(proposed) - E_initial is kept incrementally in FP64 (the accepted changes
// so far plus the proposal's), so a decision can differ from the reference's
// re-summed energies only where x lands within rounding of 1 or of U.
#include <type_traits>

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

__device__ __forceinline__ float dev_uniform(uint64_t seed, uint64_t sweep, uint32_t draw) {
  uint32_t c[4] = {(uint32_t)sweep, (uint32_t)(sweep >> 32), draw, 0x4D43u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return ((float)(c[0] >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// Reference-mode sweep in blocked rounds.  A charge's bead at its turn is its
// bead at the start of the sweep (each charge moves at most once, at its own
// turn), and the beads do not move during a sweep, so j* and the energy change
// toward each of the three possible targets (c+1, c-1, j*) against the
// start-of-sweep charge positions are computed for all charges at once
// (k_mc_prep).  The charges are then decided in blocks of MC_BLOCK: one
// wavefront walks a block in order (k_mc_round), adding the corrections for
// moves made earlier in the same block lane-parallel, and a wide pass folds the
// block's accepted moves into the energy changes of every later charge
// (k_mc_fold).  For charge k moving c -> t after charge b moved cb -> tb:
//   dE_k += q_k q_b [phi(t,tb) - phi(t,cb) - phi(c,tb) + phi(c,cb)],
// with phi(a,a) = 0: a target that coincides with a charge's bead is blocked
// at k's turn whenever that charge is still there, and otherwise the term
// cancels against the one left out before.
// The acceptance exp(-b - (E_final - E_initial)/2) is taken with
// E_final - E_initial = D + dE_k, D the sum of the accepted changes so far.
#define MC_BLOCK 64

struct McRepState {
  int cursor;       // uniforms consumed (host-fed stream)
  unsigned draws;   // device draws used
  int stop;         // stream ran out: later rounds do nothing
  int pad;
  double D;         // E_current - E_initial
};

__device__ __forceinline__ double pair_phi(const RepDev& R, const double* rx, const double* ry, const double* rz, int a,
                                           int b, int qa, int qb) {
  if (a == b) return 0.0;
  return pair_energy(R, rx, ry, rz, R.slot0 + a, R.slot0 + b, qa, qb);
}

// occupancy and start beads; state reset (one workgroup per replica)
__global__ __launch_bounds__(MC_THREADS) void k_mc_begin(const RepDev* __restrict__ reps, int rep_begin,
                                                         const int* track, int* start, unsigned char* occ,
                                                         long long occ_stride, McRepState* st, int* consumed) {
  const int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  unsigned char* oc = occ + (size_t)r * occ_stride;
  for (int i = threadIdx.x; i < R.n; i += blockDim.x) oc[i] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < R.nq; k += blockDim.x) {
    const int c = track[R.q0 + k];
    oc[c] = 1;
    start[R.q0 + k] = c;
  }
  if (threadIdx.x == 0) {
    st[r].cursor = 0;
    st[r].draws = 0;
    st[r].stop = 0;
    st[r].D = 0.0;
    if (consumed) consumed[r] = 0;
  }
}

// j* and the three candidate energy changes of each charge: one wavefront per
// charge; grid (ceil(max nq / 4), nrep)
__global__ __launch_bounds__(MC_THREADS) void k_mc_prep(const RepDev* __restrict__ reps, int rep_begin,
                                                        const double* __restrict__ rx, const double* __restrict__ ry,
                                                        const double* __restrict__ rz, const int* __restrict__ start,
                                                        const int* __restrict__ sign, int* jfirst, double* de0) {
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * (MC_THREADS / 64) + (threadIdx.x >> 6);
  if (!R.active || k >= R.nq) return;
  const int N = R.n, nq = R.nq;
  const int c = start[R.q0 + k], qk = sign[R.q0 + k];
  const int cs = R.slot0 + c;
  const double xc = rx[cs], yc = ry[cs], zc = rz[cs];
  // first j (ascending) not in {c-1,c,c+1} with |getNID(c,j)|^2 < 4.41
  int jf = -1;
  for (int base = 0; base < N; base += 64) {
    const int j = base + lane;
    bool hit = false;
    if (j < N && j != c && j != c + 1 && j != c - 1) {
      double d[3];
      const int js = R.slot0 + j;
      ref_nid(R, xc, yc, zc, rx[js], ry[js], rz[js], d[0], d[1], d[2]);
      double rr;
      {
#pragma clang fp contract(off)
        rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      }
      hit = rr < 4.41;
    }
    const unsigned long long m = __ballot(hit);
    if (m) {
      jf = base + __ffsll((long long)m) - 1;
      break;
    }
  }
  const int t0 = (c + 1 < N) ? c + 1 : -1, t1 = c - 1, t2 = jf;
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int b = lane; b < nq; b += 64) {
    if (b == k) continue;
    const int pb = start[R.q0 + b], qb = sign[R.q0 + b];
    const double ec = pair_phi(R, rx, ry, rz, c, pb, qk, qb);
    if (t0 >= 0) e0 += pair_phi(R, rx, ry, rz, t0, pb, qk, qb) - ec;
    if (t1 >= 0) e1 += pair_phi(R, rx, ry, rz, t1, pb, qk, qb) - ec;
    if (t2 >= 0) e2 += pair_phi(R, rx, ry, rz, t2, pb, qk, qb) - ec;
  }
  e0 = warp_sum(e0);
  e1 = warp_sum(e1);
  e2 = warp_sum(e2);
  if (lane == 0) {
    jfirst[R.q0 + k] = jf;
    double* d = de0 + 3 * (size_t)(R.q0 + k);
    d[0] = e0;
    d[1] = e1;
    d[2] = e2;
  }
}

// The charges k0 .. k0+MC_BLOCK-1 of each replica in order, one wavefront per
// replica.  Lane j holds charge k0+j: its start bead, sign, j*, candidate
// energy changes and the occupancy of its candidates at the start of the
// round, then its outcome.  Lane 0 reads the uniforms and writes the results.
__global__ __launch_bounds__(64) void k_mc_round(const RepDev* __restrict__ reps, int rep_begin,
                                                 const double* __restrict__ rx, const double* __restrict__ ry,
                                                 const double* __restrict__ rz, int* track, int* old,
                                                 const int* __restrict__ sign, unsigned char* occ,
                                                 long long occ_stride, const int* __restrict__ start,
                                                 const int* __restrict__ jfirst, const double* __restrict__ de0,
                                                 McRepState* st, const float* uniforms, int nu,
                                                 const uint64_t* seeds, uint64_t sweep_id, int* consumed, int* log,
                                                 const int* log_off, int* err, double* qdisp, int k0) {
  const int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active || k0 >= R.nq || st[r].stop) return;
  const int lane = threadIdx.x;
  const int N = R.n, Nb = R.nb;
  const int kend = min(k0 + MC_BLOCK, R.nq);
  unsigned char* oc = occ + (size_t)r * occ_stride;
  // this lane's charge
  const int kl = k0 + lane;
  const bool mine = kl < kend;
  int cl = 0, ql = 0, jl = -1, o0 = 1, o1 = 1, o2 = 1;
  double dl0 = 0.0, dl1 = 0.0, dl2 = 0.0;
  if (mine) {
    cl = start[R.q0 + kl];
    ql = sign[R.q0 + kl];
    jl = jfirst[R.q0 + kl];
    const double* d = de0 + 3 * (size_t)(R.q0 + kl);
    dl0 = d[0];
    dl1 = d[1];
    dl2 = d[2];
    o0 = (cl + 1 < N) ? oc[cl + 1] : 1;
    o1 = (cl >= 1) ? oc[cl - 1] : 1;
    o2 = (jl >= 0) ? oc[jl] : 1;
  }
  int cursor = st[r].cursor;
  unsigned draws = st[r].draws;
  double D = st[r].D;
  const float* ur = uniforms ? uniforms + (size_t)r * nu : nullptr;
  // the round's uniforms (at most two per charge)
  float uw0 = 2.0f, uw1 = 2.0f;
  const int cur0 = cursor;
  if (ur) {
    if (cur0 + lane < nu) uw0 = ur[cur0 + lane];
    if (cur0 + 64 + lane < nu) uw1 = ur[cur0 + 64 + lane];
  }
  bool moved = false;
  int tb = -1;
  bool stop = false;
  for (int k = k0; k < kend; ++k) {
    const int i = k - k0;
    const int c = __shfl(cl, i, 64), qk = __shfl(ql, i, 64), jf = __shfl(jl, i, 64);
    // draw mm
    float mmf;
    if (ur) {
      if (cursor >= nu) {
        stop = true;
        break;
      }
      const int w = cursor - cur0;
      mmf = (w < 64) ? __shfl(uw0, w, 64) : __shfl(uw1, w - 64, 64);
      ++cursor;
    } else {
      mmf = dev_uniform(seeds[r], sweep_id, draws++);
    }
    const double mm = (double)mmf;
    int target, kind, a;
    if (jf >= 0) {
      if (mm < .33) {
        target = c + 1;
        kind = 1;
        a = 0;
      } else if (mm > 0.66) {
        target = c - 1;
        kind = 1;
        a = 1;
      } else {
        target = jf;
        kind = 2;
        a = 2;
      }
    } else {
      kind = 0;
      if (mm > .5) {
        target = c + 1;
        a = 0;
      } else {
        target = c - 1;
        a = 1;
      }
    }
    // occupancy of the target now: at the start of the round, less a charge
    // of this block that left it, plus one that moved onto it
    const int oa = (a == 0) ? o0 : (a == 1) ? o1 : o2;
    const bool occ0 = __shfl(oa, i, 64) != 0;
    const bool freed = __ballot(moved && cl == target) != 0ull;
    const bool filled = __ballot(moved && tb == target) != 0ull;
    const bool occupied = (occ0 && !freed) || filled;
    bool blocked;
    if (a == 0)
      blocked = (c == N - 1) || (c % Nb == Nb - 1) || (c == Nb - 1) || occupied;
    else if (a == 1)
      blocked = (c - 1 < 0) || (c == 0) || (c % Nb == 0) || occupied;
    else
      blocked = occupied;
    int* lg = log ? log + 6 * (size_t)(log_off[r] + k) : nullptr;
    if (blocked) {
      if (lane == 0 && lg) {
        lg[0] = k;
        lg[1] = c;
        lg[2] = -1;
        lg[3] = -1;
        lg[4] = 0;
        lg[5] = 1;
      }
      continue;
    }
    // energy change: the precomputed value plus this block's earlier moves
    double corr = 0.0;
    if (moved) {
      const int qb = ql;
      corr = pair_phi(R, rx, ry, rz, target, tb, qk, qb) - pair_phi(R, rx, ry, rz, target, cl, qk, qb) -
             pair_phi(R, rx, ry, rz, c, tb, qk, qb) + pair_phi(R, rx, ry, rz, c, cl, qk, qb);
    }
    corr = warp_sum(corr);
    const double dsel = (a == 0) ? dl0 : (a == 1) ? dl1 : dl2;
    const double dE = __shfl(dsel, i, 64) + corr;
    const double bar = (kind == 2) ? R.barrier2 : R.barrier;
    const double x = exp((-bar - 0.5 * (D + dE)));
    int ndraw = 1, acc = 0;
    if (1 <= x) {
      acc = 1;
    } else {
      float dice;
      bool ok = true;
      if (ur) {
        if (cursor >= nu) {
          ok = false;
          dice = 2.0f;
        } else {
          const int w = cursor - cur0;
          dice = (w < 64) ? __shfl(uw0, w, 64) : __shfl(uw1, w - 64, 64);
          ++cursor;
        }
      } else {
        dice = dev_uniform(seeds[r], sweep_id, draws++);
      }
      ndraw = 2;
      if (!ok && lane == 0) atomicOr(err, 2);
      if (ok && (double)dice <= x) acc = 1;
    }
    if (acc) D += dE;
    if (lane == i) {
      moved = acc != 0;
      tb = target;
    }
    if (lane == 0) {
      if (acc) {
        track[R.q0 + k] = target;
        old[R.q0 + k] = c;
        oc[c] = 0;
        oc[target] = 1;
        if (qdisp) {  // the hop joins the charge's displacement log (dxt of chargeHop)
          double d[3];
          hop_vector(R, rx, ry, rz, R.slot0 + target, R.slot0 + c, d);
          for (int q = 0; q < 3; ++q) qdisp[3 * (R.q0 + k) + q] += d[q];
        }
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
  }
  if (lane == 0) {
    if (stop) {
      atomicOr(err, 2);
      st[r].stop = 1;
    }
    st[r].cursor = cursor;
    st[r].draws = draws;
    st[r].D = D;
    if (consumed) consumed[r] = ur ? cursor : (int)draws;
  }
}

// The moves of block k0 .. k0+MC_BLOCK-1 folded into the candidate energy
// changes of every later charge: one wavefront per later charge, lane j for
// charge k0+j of the block; grid (ceil(max nq / 4), nrep)
__global__ __launch_bounds__(MC_THREADS) void k_mc_fold(const RepDev* __restrict__ reps, int rep_begin,
                                                        const double* __restrict__ rx, const double* __restrict__ ry,
                                                        const double* __restrict__ rz, const int* __restrict__ track,
                                                        const int* __restrict__ start, const int* __restrict__ sign,
                                                        const int* __restrict__ jfirst, const McRepState* st,
                                                        double* de0, int k0) {
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * (MC_THREADS / 64) + (threadIdx.x >> 6);
  const int kend = min(k0 + MC_BLOCK, R.nq);
  if (!R.active || k < kend || k >= R.nq || st[r].stop) return;
  const int b = k0 + lane;
  int cb = -1, tb = -1, qb = 0;
  bool moved = false;
  if (b < kend) {
    cb = start[R.q0 + b];
    tb = track[R.q0 + b];
    qb = sign[R.q0 + b];
    moved = tb != cb;
  }
  if (__ballot(moved) == 0ull) return;
  const int N = R.n;
  const int c = start[R.q0 + k], qk = sign[R.q0 + k], jf = jfirst[R.q0 + k];
  const int t0 = (c + 1 < N) ? c + 1 : -1, t1 = c - 1, t2 = jf;
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  if (moved) {
    const double ec = pair_phi(R, rx, ry, rz, c, tb, qk, qb) - pair_phi(R, rx, ry, rz, c, cb, qk, qb);
    if (t0 >= 0) e0 = pair_phi(R, rx, ry, rz, t0, tb, qk, qb) - pair_phi(R, rx, ry, rz, t0, cb, qk, qb) - ec;
    if (t1 >= 0) e1 = pair_phi(R, rx, ry, rz, t1, tb, qk, qb) - pair_phi(R, rx, ry, rz, t1, cb, qk, qb) - ec;
    if (t2 >= 0) e2 = pair_phi(R, rx, ry, rz, t2, tb, qk, qb) - pair_phi(R, rx, ry, rz, t2, cb, qk, qb) - ec;
  }
  e0 = warp_sum(e0);
  e1 = warp_sum(e1);
  e2 = warp_sum(e2);
  if (lane == 0) {
    double* d = de0 + 3 * (size_t)(R.q0 + k);
    d[0] += e0;
    d[1] += e1;
    d[2] += e2;
  }
}

size_t mc_state_bytes() { return sizeof(McRepState); }

void launch_mc_sweep(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_nq, const double* rx,
                     const double* ry, const double* rz, int* track, int* old, const int* sign, unsigned char* occ,
                     long long occ_stride, int* start, int* jfirst, double* de0, void* state, const float* uniforms,
                     int nu, const uint64_t* seeds, uint64_t sweep_id, int* consumed, int* log, const int* log_off,
                     int* err, double* qdisp) {
  if (nrep <= 0) return;
  McRepState* S = (McRepState*)state;
  k_mc_begin<<<nrep, MC_THREADS, 0, st>>>(reps, rep_begin, track, start, occ, occ_stride, S, consumed);
  if (max_nq <= 0) return;
  const dim3 wide((max_nq + MC_THREADS / 64 - 1) / (MC_THREADS / 64), nrep);
  k_mc_prep<<<wide, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, start, sign, jfirst, de0);
  for (int k0 = 0; k0 < max_nq; k0 += MC_BLOCK) {
    k_mc_round<<<nrep, 64, 0, st>>>(reps, rep_begin, rx, ry, rz, track, old, sign, occ, occ_stride, start, jfirst, de0,
                                    S, uniforms, nu, seeds, sweep_id, consumed, log, log_off, err, qdisp, k0);
    if (k0 + MC_BLOCK < max_nq)
      k_mc_fold<<<wide, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, track, start, sign, jfirst, S, de0, k0);
  }
}

// Energy of all charge pairs of each replica (the double loop of chargeHop,
// S/main26.c:1955-1970, over the charged beads): one wavefront per charge a
// sums its pairs with b > a, then one workgroup per replica adds the rows in
// order.
__global__ __launch_bounds__(MC_THREADS) void k_mc_energy_rows(const RepDev* __restrict__ reps, int rep_begin,
                                                               const double* __restrict__ rx,
                                                               const double* __restrict__ ry,
                                                               const double* __restrict__ rz,
                                                               const int* __restrict__ track,
                                                               const int* __restrict__ sign, double* rows) {
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int lane = threadIdx.x & 63;
  const int a = blockIdx.x * (MC_THREADS / 64) + (threadIdx.x >> 6);
  if (!R.active || a >= R.nq) return;
  const int as = R.slot0 + track[R.q0 + a], qa = sign[R.q0 + a];
  double e = 0.0;
  for (int b = a + 1 + lane; b < R.nq; b += 64)
    e += pair_energy(R, rx, ry, rz, as, R.slot0 + track[R.q0 + b], qa, sign[R.q0 + b]);
  e = warp_sum(e);
  if (lane == 0) rows[R.q0 + a] = e;
}

__global__ __launch_bounds__(MC_THREADS) void k_mc_energy_sum(const RepDev* __restrict__ reps, int rep_begin,
                                                              const double* __restrict__ rows, double* e) {
  const int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  __shared__ double sh[MC_THREADS / 64];
  double v = 0.0;
  for (int a = threadIdx.x; a < R.nq; a += blockDim.x) v += rows[R.q0 + a];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) e[r] = v;
}

void launch_mc_energy(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_nq, const double* rx,
                      const double* ry, const double* rz, const int* track, const int* sign, double* rows,
                      double* e) {
  if (nrep <= 0) return;
  if (max_nq > 0) {
    const dim3 g((max_nq + MC_THREADS / 64 - 1) / (MC_THREADS / 64), nrep);
    k_mc_energy_rows<<<g, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, track, sign, rows);
  }
  k_mc_energy_sum<<<nrep, MC_THREADS, 0, st>>>(reps, rep_begin, rows, e);
}

// ============================================================================
// Exactly detailed-balanced hops (no reference counterpart; same energy and
// barriers as chargeHop).  Attempt a: u0 picks charge k (at bead x), u1 picks a
// destination y among the uncharged beads within the hop radius of x by the
// inverse CDF of q(x->.) in ascending bead order, u2 is the acceptance dice:
//   rho = exp(-dE) q(y->x) / q(x->y),  a = exp(-b) (b = barrier for a chain
//   neighbour, barrier2 otherwise),  A = a sqrt(rho) / max(1, a sqrt(rho), a/sqrt(rho)).
// q is uniform, or (learned) (1-eps) w_c / sum(w) + eps/ncand with w_c =
// sigmoid(MLP score), the network (trained to predict acceptance of uniform
// proposals) scoring 16 features of each candidate move (BF16 weights, FP32
// arithmetic).
// q(y->x) is evaluated in the state after the move.  The beads within
// 2 hop + feature radius of x (all any forward or reverse feature can see),
// ascending, come from per-bead lists built once per sweep (the beads do not
// move during a sweep), each entry flagged when it lies within the hop radius;
// features are summed by one wavefront per candidate.  dE is exact FP64 over
// all charges: with V(z) = sum_b q_b phi(z, b) over every charge b (phi(z,z) = 0)
//   dE = lambda_B q_k [V(y) - V(x)] - lambda_B phi(x, y),
// V computed for every bead at the start of the sweep and brought up to date
// after each round of MH_ROUND attempts by a wide pass over the round's
// accepted moves; within a round the earlier moves' changes of V at x and y
// are added one term per thread.
#define MH_MAXL 4096
#define MH_MAXC 128
#define MLP_NP (16 * 64 + 64 + 64 * 64 + 64 + 64 + 1)

#define MH_ROUND 64
// workgroup sizes of the MH round: the learned proposal scores its candidates one
// wavefront each, so it takes more wavefronts than the uniform one
#define MH_THREADS_UNIFORM MC_THREADS
#define MH_THREADS_LEARNED 512
#define MH_THREADS_MAX 1024
#define MH_HOP_PRE 16  // hop-list entries of each attempt's bead fetched with the round
// replicas up to these sizes keep the occupancy and charge beads in LDS
#define MH_LDS_N 16384
#define MH_LDS_Q 4096
// neighbourhoods up to this size keep the source-side distances of the
// features in LDS
#define MH_SRC_L 1536
struct MhState {
  signed char qs[MH_LDS_N];
  int trk[MH_LDS_Q];
};
struct MhNoState {
  int unused;
};

struct MhShared {
  int loc[MH_MAXL];
  double vx[MH_SRC_L], vy[MH_SRC_L], vz[MH_SRC_L];  // loc[] relative to the attempt's start bead
  double rs[MH_SRC_L], es[MH_SRC_L];                // distance to the proposal's source, exp(-kappa r)/r
  int4 mv[MH_ROUND];
  float ur[3 * MH_ROUND];
  int xp[MH_ROUND], op[MH_ROUND], oe[MH_ROUND];  // the round's attempts: bead, list ranges
  int hp[MH_ROUND], he[MH_ROUND];                 // and V at the start of the round,
  double vxp[MH_ROUND];                           // the charge's sign and x's hop list
  int qkp[MH_ROUND];
  int hl[MH_ROUND][MH_HOP_PRE];
  int nmv;
  int cand[MH_MAXC];
  double q[MH_MAXC];
  double gdst[MH_MAXC][5];  // learned proposal: each candidate's geometry (mh_geom_dst)
  double gsrc[4];           // and the source's (mh_geom_src)
  float feat[MH_MAXC][16];
  float fsel[16];
  float w[MLP_NP];
  double dsum[MH_THREADS_MAX / 64];
  int wcnt[MH_THREADS_MAX / 64];
  int nloc, ncand, k, x, qk, y, pick, accept, bad;
  float u[3];
  double qf, zq;
};

__device__ __forceinline__ double mh_dist(const RepDev& R, const double* rx, const double* ry, const double* rz, int a,
                                          int b) {
  const int sa = R.slot0 + a, sb = R.slot0 + b;
  double d[3];
  ref_nid(R, rx[sa], ry[sa], rz[sa], rx[sb], ry[sb], rz[sb], d[0], d[1], d[2]);
  return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

__device__ __forceinline__ double mh_dist2(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                           int a, int b) {
  const int sa = R.slot0 + a, sb = R.slot0 + b;
  double d[3];
  ref_nid(R, rx[sa], ry[sa], rz[sa], rx[sb], ry[sb], rz[sb], d[0], d[1], d[2]);
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

__device__ __forceinline__ double mh_cos_at(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                            int b) {
  const int pos = b % R.nb;
  if (pos == 0 || pos == R.nb - 1) return 0.0;
  const int s = R.slot0 + b;
  double u[3], v[3];
  ref_nid(R, rx[s - 1], ry[s - 1], rz[s - 1], rx[s], ry[s], rz[s], u[0], u[1], u[2]);
  ref_nid(R, rx[s + 1], ry[s + 1], rz[s + 1], rx[s], ry[s], rz[s], v[0], v[1], v[2]);
  double uu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), vv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) / (uu * vv);
}

// Append the flagged values of one block-wide round in ascending thread order.
__device__ __forceinline__ void block_append(bool flag, int value, int* list, int* cnt, int cap, int* wcnt) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long m = __ballot(flag);
  if (lane == 0) wcnt[w] = __popcll(m);
  __syncthreads();
  int off = *cnt;
  for (int k = 0; k < w; ++k) off += wcnt[k];
  int pos = off + __popcll(m & ((1ull << lane) - 1ull));
  if (flag && pos < cap) list[pos] = value;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += wcnt[k];
    *cnt += t;
  }
  __syncthreads();
}

// Distance between two beads from their vectors relative to a common anchor.
__device__ __forceinline__ double mh_rel_dist(double ax, double ay, double az, double bx, double by, double bz) {
#pragma clang fp contract(off)
  const double dx = bx - ax, dy = by - ay, dz = bz - az;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// 16 features of moving the charge at src (sign qk) to dst, summed by one
// wavefront over the gathered neighbourhood; clr / set (or -1) give the state
// after a move.  The distances of src and dst to the neighbourhood beads are
// taken from the beads' getNID vectors relative to the anchor (the attempt's
// start bead): vx/vy/vz hold them for loc[] (null: computed here); rs / es
// (may be null) hold the distance to src and exp(-kappa r)/r at it.  Lane 0
// writes out[16].
__device__ void mh_features_rel(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                const signed char* qb, const int* loc, int nloc, int src, int dst, int qk, int clr,
                                int set, int ncand, float* out, int anchor, const double* vx, const double* vy,
                                const double* vz, const double* rs, const double* es, const double* geo_src = nullptr,
                                const double* geo_dst = nullptr) {
  const int lane = threadIdx.x & 63;
  double e = 0, n4d = 0, n4s = 0, s4d = 0, s4s = 0, b4d = 0, b4s = 0;
  const double R4 = R.feat_r, kap = R.kappa_d, R4hi = R4 * R4 * (1.0 + 1e-9);
  const int sa = R.slot0 + anchor, ss = R.slot0 + src, sd = R.slot0 + dst;
  double vs[3], vd[3], f0, f10, f11;
  if (geo_src && geo_dst) {  // precomputed by mh_geom_src / mh_geom_dst (the same arithmetic)
    for (int d = 0; d < 3; ++d) {
      vs[d] = geo_src[d];
      vd[d] = geo_dst[d];
    }
    f11 = geo_src[3];
    f0 = geo_dst[3];
    f10 = geo_dst[4];
  } else {
    ref_nid(R, rx[ss], ry[ss], rz[ss], rx[sa], ry[sa], rz[sa], vs[0], vs[1], vs[2]);
    ref_nid(R, rx[sd], ry[sd], rz[sd], rx[sa], ry[sa], rz[sa], vd[0], vd[1], vd[2]);
    // the pair's own geometry, ahead of the loop so that its loads overlap the above
    f0 = mh_dist(R, rx, ry, rz, src, dst);
    f10 = mh_cos_at(R, rx, ry, rz, dst);
    f11 = mh_cos_at(R, rx, ry, rz, src);
  }
  for (int i = lane; i < nloc; i += 64) {
    const int m = loc[i];
    if (m == src || m == dst) continue;
    double vm[3];
    if (vx) {
      vm[0] = vx[i];
      vm[1] = vy[i];
      vm[2] = vz[i];
    } else {
      const int sm = R.slot0 + m;
      ref_nid(R, rx[sm], ry[sm], rz[sm], rx[sa], ry[sa], rz[sa], vm[0], vm[1], vm[2]);
    }
    // a bead beyond the feature radius of both src and dst adds nothing; the
    // squared distances decide that without the square roots (R4hi lies above
    // R4^2 by more than the rounding of the root)
    double dd2, ds2 = 0.0;
    {
#pragma clang fp contract(off)
      const double dx = vm[0] - vd[0], dy = vm[1] - vd[1], dz = vm[2] - vd[2];
      dd2 = dx * dx + dy * dy + dz * dz;
      if (!rs) {
        const double sx = vm[0] - vs[0], sy = vm[1] - vs[1], sz = vm[2] - vs[2];
        ds2 = sx * sx + sy * sy + sz * sz;
      }
    }
    const bool near_d = dd2 < R4hi;
    const bool near_s = rs ? rs[i] < R4 : ds2 < R4hi;
    if (!near_d && !near_s) continue;
    const double rdm = sqrt(dd2);
    const double rsm = rs ? rs[i] : sqrt(ds2);
    const int qm = (m == clr) ? 0 : (m == set) ? qk : (int)qb[m];
    if (qm != 0 && (rdm < R4 || rsm < R4))
      e += R.lambda_b * (double)(qk * qm) * (exp(-kap * rdm) / rdm - (es ? es[i] : exp(-kap * rsm) / rsm));
    if (rdm < R4) {
      b4d += 0.1;
      if (qm != 0) {
        n4d += 1.0;
        s4d += (double)(qk * qm);
      }
    }
    if (rsm < R4) {
      b4s += 0.1;
      if (qm != 0) {
        n4s += 1.0;
        s4s += (double)(qk * qm);
      }
    }
  }
  e = warp_sum(e);
  n4d = warp_sum(n4d);
  n4s = warp_sum(n4s);
  s4d = warp_sum(s4d);
  s4s = warp_sum(s4s);
  b4d = warp_sum(b4d);
  b4s = warp_sum(b4s);
  if (lane == 0) {
    const int nb = R.nb, gap = abs(dst - src);
    const bool same = (src / nb) == (dst / nb);
    out[0] = (float)f0;
    out[1] = (same && gap == 1) ? 1.f : 0.f;
    out[2] = same ? (float)((double)(gap < 20 ? gap : 20) / 20.0) : 1.f;
    out[3] = (float)e;
    out[4] = (float)n4d;
    out[5] = (float)n4s;
    out[6] = (float)s4d;
    out[7] = (float)s4s;
    out[8] = (float)b4d;
    out[9] = (float)b4s;
    out[10] = (float)f10;
    out[11] = (float)f11;
    out[12] = (dst % nb == 0 || dst % nb == nb - 1) ? 1.f : 0.f;
    out[13] = (src % nb == 0 || src % nb == nb - 1) ? 1.f : 0.f;
    out[14] = (float)qk;
    out[15] = (float)((double)ncand / 10.0);
  }
}

// The source side of a proposal's features: src relative to the anchor and the
// bend cosine at src, as mh_features_rel computes them.
__device__ __forceinline__ void mh_geom_src(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                            int src, int anchor, double* g) {
  const int sa = R.slot0 + anchor, ss = R.slot0 + src;
  ref_nid(R, rx[ss], ry[ss], rz[ss], rx[sa], ry[sa], rz[sa], g[0], g[1], g[2]);
  g[3] = mh_cos_at(R, rx, ry, rz, src);
}

// One candidate's side: dst relative to the anchor, |src - dst| and the bend
// cosine at dst, as mh_features_rel computes them.
__device__ __forceinline__ void mh_geom_dst(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                            int src, int dst, int anchor, double* g) {
  const int sa = R.slot0 + anchor, sd = R.slot0 + dst;
  ref_nid(R, rx[sd], ry[sd], rz[sd], rx[sa], ry[sa], rz[sa], g[0], g[1], g[2]);
  g[3] = mh_dist(R, rx, ry, rz, src, dst);
  g[4] = mh_cos_at(R, rx, ry, rz, dst);
}

// The network's weights as the MH kernels keep them in LDS: W1 and W2
// transposed (W1T[i][j] = W1[j][i], W2T[i][j] = W2[j][i]) so that lane j's reads
// of its hidden unit's row fall on consecutive banks; then b1, b2, W3, b3.
__device__ __forceinline__ int mlp_lds_index(int g) {
  if (g < 64 * 16) return (g % 16) * 64 + g / 16;  // W1[j][i] -> W1T[i][j]
  g -= 64 * 16;
  if (g < 64) return 64 * 16 + 64 * 64 + g;  // b1 after both matrices
  g -= 64;
  if (g < 64 * 64) return 64 * 16 + (g % 64) * 64 + g / 64;  // W2[j][i] -> W2T[i][j]
  return 64 * 16 + 64 * 64 + 64 + (g - 64 * 64);  // b2, W3, b3
}

// MLP score of one feature vector by one wavefront: lane j owns hidden unit j;
// the first layer's outputs reach every lane by readlane, in order.
__device__ float mlp_score_wave(const float* w, const float* f) {
  const int j = threadIdx.x & 63;
  const float *W1T = w, *W2T = W1T + 64 * 16, *b1 = W2T + 64 * 64, *b2 = b1 + 64, *W3 = b2 + 64, *b3 = W3 + 64;
  float a = b1[j];
#pragma unroll
  for (int i = 0; i < 16; ++i) a = fmaf(W1T[i * 64 + j], f[i], a);
  const float h1 = fmaxf(a, 0.f);
  float a2 = b2[j];
  for (int i = 0; i < 64; ++i) a2 = fmaf(W2T[i * 64 + j], __shfl(h1, i, 64), a2);
  const float h2 = fmaxf(a2, 0.f);
  float s = W3[j] * h2;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s + b3[0];
}

// proposal over the scored candidates (wavefront 0; the caller synchronises before
// and after): the network's predicted acceptance sigmoid(score) normalised (a
// Barker-type locally balanced weight), mixed with the uniform proposal at weight
// eps.  The elementwise parts run lane-parallel and the normaliser is summed by
// lane 0 in candidate order, so every value equals a sequential evaluation.
__device__ void mh_proposal_probs(double* q, int nc, double eps, double* zsh) {
  const int lane = threadIdx.x & 63;
  for (int c = lane; c < nc; c += 64) q[c] = 1.0 / (1.0 + exp(-q[c]));
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane == 0) {
    double z = 0.0;
    for (int c = 0; c < nc; ++c) z += q[c];
    *zsh = z;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double z = *zsh;
  for (int c = lane; c < nc; c += 64) q[c] = (1.0 - eps) * q[c] / z + eps / nc;
}

__device__ __forceinline__ double mh_phi(const RepDev& R, const double* rx, const double* ry, const double* rz, int a,
                                         int b) {
  if (a == b) return 0.0;
  const double r = mh_dist(R, rx, ry, rz, a, b);
  return exp(-R.kappa_d * r) / r;
}

// per-bead lists, ascending: the neighbourhood (within 2 hop + feature radius)
// and the hop list (within the hop radius); one wavefront per bead, grid
// (ceil(max n / 4), nrep)
__global__ __launch_bounds__(MC_THREADS) void k_mh_count(const RepDev* __restrict__ reps, int rep_begin,
                                                         const double* __restrict__ rx, const double* __restrict__ ry,
                                                         const double* __restrict__ rz, int* cnt, int* hcnt) {
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (MC_THREADS / 64) + (threadIdx.x >> 6);
  if (!R.active || R.nq == 0 || i >= R.n) return;
  int c = 0, h = 0;
  for (int base = 0; base < R.n; base += 64) {
    const int j = base + lane;
    double d2 = 0.0;
    if (j < R.n) d2 = mh_dist2(R, rx, ry, rz, i, j);
    c += __popcll(__ballot(j < R.n && d2 < R.nbhd2));
    h += __popcll(__ballot(j < R.n && d2 < R.hop2));
  }
  if (lane == 0) {
    cnt[R.slot0 + i] = c;
    hcnt[R.slot0 + i] = h;
  }
}

__global__ __launch_bounds__(MC_THREADS) void k_mh_fill(const RepDev* __restrict__ reps, int rep_begin,
                                                        const double* __restrict__ rx, const double* __restrict__ ry,
                                                        const double* __restrict__ rz, const int* __restrict__ off,
                                                        int* list, const int* __restrict__ hoff, int* hlist) {
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (MC_THREADS / 64) + (threadIdx.x >> 6);
  if (!R.active || R.nq == 0 || i >= R.n) return;
  int* out = list + off[R.slot0 + i];
  int* hout = hlist + hoff[R.slot0 + i];
  int pos = 0, hpos = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int base = 0; base < R.n; base += 64) {
    const int j = base + lane;
    double d2 = 0.0;
    if (j < R.n) d2 = mh_dist2(R, rx, ry, rz, i, j);
    const bool f = j < R.n && d2 < R.nbhd2, h = j < R.n && d2 < R.hop2;
    const unsigned long long m = __ballot(f), mh = __ballot(h);
    if (f) out[pos + __popcll(m & below)] = j;
    if (h) hout[hpos + __popcll(mh & below)] = j;
    pos += __popcll(m);
    hpos += __popcll(mh);
  }
}

// V(z) for every bead of a replica with charges: one wavefront per bead
__global__ __launch_bounds__(MC_THREADS) void k_mh_potential(const RepDev* __restrict__ reps, int rep_begin,
                                                             const double* __restrict__ rx,
                                                             const double* __restrict__ ry,
                                                             const double* __restrict__ rz,
                                                             const int* __restrict__ track,
                                                             const int* __restrict__ sign, double* V) {
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int lane = threadIdx.x & 63;
  const int z = blockIdx.x * (MC_THREADS / 64) + (threadIdx.x >> 6);
  if (!R.active || R.nq == 0 || z >= R.n) return;
  double v = 0.0;
  for (int b = lane; b < R.nq; b += 64) v += (double)sign[R.q0 + b] * mh_phi(R, rx, ry, rz, z, track[R.q0 + b]);
  v = warp_sum(v);
  if (lane == 0) V[R.slot0 + z] = v;
}

// occupancy (signed charge per bead) and the per-replica counters
__global__ __launch_bounds__(MC_THREADS) void k_mh_begin(const RepDev* __restrict__ reps, int rep_begin,
                                                         const int* track, const int* sign, signed char* occ,
                                                         long long occ_stride, int* done, int* stop, int* consumed) {
  const int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  signed char* qb = occ + (size_t)r * occ_stride;
  for (int i = threadIdx.x; i < R.n; i += blockDim.x) qb[i] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < R.nq; k += blockDim.x) qb[track[R.q0 + k]] = (signed char)sign[R.q0 + k];
  if (threadIdx.x == 0) {
    done[r] = 0;
    stop[r] = 0;
    if (consumed) consumed[r] = 0;
  }
}

// The changes of V by a round's accepted moves: one wavefront per bead, lane m
// for move m
__global__ __launch_bounds__(MC_THREADS) void k_mh_fold(const RepDev* __restrict__ reps, int rep_begin,
                                                        const double* __restrict__ rx, const double* __restrict__ ry,
                                                        const double* __restrict__ rz, const int4* __restrict__ moves,
                                                        const int* __restrict__ nmoves, double* V) {
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int lane = threadIdx.x & 63;
  const int z = blockIdx.x * (MC_THREADS / 64) + (threadIdx.x >> 6);
  if (!R.active || z >= R.n) return;
  const int nm = nmoves[r];
  if (nm == 0) return;
  double v = 0.0;
  if (lane < nm) {
    const int4 mv = moves[(size_t)r * MH_ROUND + lane];
    v = (double)mv.w * (mh_phi(R, rx, ry, rz, z, mv.z) - mh_phi(R, rx, ry, rz, z, mv.y));
  }
  v = warp_sum(v);
  if (lane == 0) V[R.slot0 + z] += v;
}

// Attempts a0 .. a0 + MH_ROUND - 1 of each replica (one workgroup per replica).
// LQ: the occupancy and the charges' beads are read from LDS copies (written
// through to memory on every accepted move).
template <bool LQ, int NT>
__global__ __launch_bounds__(NT) void k_mh_round(
    const RepDev* __restrict__ reps, int rep_begin, const double* __restrict__ rx, const double* __restrict__ ry,
    const double* __restrict__ rz, int* track, int* old, const int* sign, signed char* occ, long long occ_stride,
    const float* uniforms, int nu, const uint64_t* seeds, uint64_t sweep_id, const float* weights, int* consumed,
    int* mlog, float* feats, const int* log_off, int* err, double* qdisp, const int* __restrict__ nb_off,
    const int* __restrict__ nb_list, const int* __restrict__ hop_off, const int* __restrict__ hop_list,
    const double* __restrict__ V, int4* moves, int* nmoves, int* done, int* stop, int a0) {
  const int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  const int nq = R.nq, tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  if (a0 >= nq || stop[r]) {
    if (tid == 0) nmoves[r] = 0;
    return;
  }
  __shared__ MhShared S;
  __shared__ typename std::conditional<LQ, MhState, MhNoState>::type L;
  const int nwv = blockDim.x >> 6;
  const bool learned = weights != nullptr;
  signed char* qg = occ + (size_t)r * occ_stride;
  int* tg = track + R.q0;
  signed char* qb = qg;
  int* tv = tg;
  if constexpr (LQ) {
    qb = L.qs;
    tv = L.trk;
    for (int i = tid; i < R.n; i += blockDim.x) qb[i] = qg[i];
    for (int i = tid; i < nq; i += blockDim.x) tv[i] = tg[i];
  }
  if (learned)
    for (int i = tid; i < MLP_NP; i += blockDim.x) S.w[mlp_lds_index(i)] = weights[i];
  const int aend = min(a0 + MH_ROUND, nq);
  const float* ur = uniforms ? uniforms + (size_t)r * nu : nullptr;
  for (int i = tid; i < 3 * (aend - a0); i += blockDim.x) {
    const int g = 3 * a0 + i;
    S.ur[i] = ur ? (g < nu ? ur[g] : -1.0f) : dev_uniform(seeds[r], sweep_id, (uint32_t)g);
  }
  if (tid == 0) S.nmv = 0;
  __syncthreads();
  // each attempt's bead as of the start of the round, its list range and V there
  for (int i = tid; i < aend - a0; i += blockDim.x) {
    const float u0 = S.ur[3 * i];
    int kk = (int)((u0 < 0.0f ? 0.0f : u0) * (float)nq);
    kk = kk > nq - 1 ? nq - 1 : kk;
    const int xp = tv[kk];
    S.xp[i] = xp;
    S.qkp[i] = sign[R.q0 + kk];
    S.op[i] = nb_off[R.slot0 + xp];
    S.oe[i] = nb_off[R.slot0 + xp + 1];
    const int h0 = hop_off[R.slot0 + xp], h1 = hop_off[R.slot0 + xp + 1];
    S.hp[i] = h0;
    S.he[i] = h1;
    S.vxp[i] = V[R.slot0 + xp];
    for (int e = 0; e < MH_HOP_PRE && e < h1 - h0; ++e) S.hl[i][e] = hop_list[h0 + e];
  }
  __syncthreads();
  int done_att = done[r];
  bool halted = false;
  for (int a = a0; a < aend; ++a) {
    if (tid == 0) {
      S.bad = 0;
      for (int i = 0; i < 3; ++i) {
        const float u = S.ur[3 * (a - a0) + i];
        if (u < 0.0f) S.bad = 1;
        S.u[i] = u < 0.0f ? 0.f : u;
      }
      int k = (int)(S.u[0] * (float)nq);
      S.k = k > nq - 1 ? nq - 1 : k;
      S.x = tv[S.k];
      S.qk = S.qkp[a - a0];
      S.ncand = 0;
    }
    __syncthreads();
    if (S.bad) {
      if (tid == 0) atomicOr(err, 2);
      halted = true;
      break;
    }
    const int k = S.k, x = S.x, qk = S.qk;
    int* lg = mlog ? mlog + 6 * (size_t)(log_off[r] + a) : nullptr;
    float* fa = feats ? feats + 16 * (size_t)(log_off[r] + a) : nullptr;
    // neighbourhood of x (x included), ascending, from its list: needed by the
    // features only
    const bool need_loc = learned || fa != nullptr;
    const bool pre = S.xp[a - a0] == x;
    const int o0 = pre ? S.op[a - a0] : nb_off[R.slot0 + x];
    const int nloc = (pre ? S.oe[a - a0] : nb_off[R.slot0 + x + 1]) - o0;
    const int h0 = pre ? S.hp[a - a0] : hop_off[R.slot0 + x];
    const int nh = (pre ? S.he[a - a0] : hop_off[R.slot0 + x + 1]) - h0;
    const double vx = pre ? S.vxp[a - a0] : V[R.slot0 + x];
    if (nloc > MH_MAXL) {
      if (tid == 0) atomicOr(err, 8);
      halted = true;
      break;
    }
    if (need_loc)
      for (int i = tid; i < nloc; i += blockDim.x) S.loc[i] = nb_list[o0 + i];
    // candidates of x: its hop list less x and the charged beads (one wavefront)
    if (wv == 0) {
      int cnt = 0;
      for (int base = 0; base < nh; base += 64) {
        const int i = base + lane;
        const int b = i < nh ? ((pre && nh <= MH_HOP_PRE) ? S.hl[a - a0][i] : hop_list[h0 + i]) : -1;
        const bool f = b >= 0 && b != x && qb[b] == 0;
        const unsigned long long m = __ballot(f);
        const int p = cnt + __popcll(m & ((1ull << lane) - 1ull));
        if (f && p < MH_MAXC) S.cand[p] = b;
        cnt += __popcll(m);
      }
      if (lane == 0) S.ncand = cnt;
    }
    __syncthreads();
    const int nc = S.ncand;
    if (nc > MH_MAXC) {
      if (tid == 0) atomicOr(err, 8);
      halted = true;
      break;
    }
    const bool src_lds = nloc <= MH_SRC_L;
    if (need_loc && src_lds && nc > 0) {
      const int sx = R.slot0 + x;
      for (int i = tid; i < nloc; i += blockDim.x) {
        const int sm = R.slot0 + S.loc[i];
        ref_nid(R, rx[sm], ry[sm], rz[sm], rx[sx], ry[sx], rz[sx], S.vx[i], S.vy[i], S.vz[i]);
      }
      __syncthreads();
    }
    const double* VX = src_lds ? S.vx : nullptr;
    const double* VY = src_lds ? S.vy : nullptr;
    const double* VZ = src_lds ? S.vz : nullptr;
    if (nc == 0) {
      if (tid == 0) {
        if (lg) {
          lg[0] = k;
          lg[1] = x;
          lg[2] = -1;
          lg[3] = -1;
          lg[4] = 0;
          lg[5] = 3;
        }
        if (fa)
          for (int i = 0; i < 16; ++i) fa[i] = 0.f;
      }
      done_att = a + 1;
      __syncthreads();
      continue;
    }
    // forward proposal
    if (learned) {
      if (src_lds) {  // x's own vector is zero
        for (int i = tid; i < nloc; i += blockDim.x) {
          if (S.loc[i] == x) continue;
          const double rsm = mh_rel_dist(0.0, 0.0, 0.0, S.vx[i], S.vy[i], S.vz[i]);
          S.rs[i] = rsm;
          S.es[i] = exp(-R.kappa_d * rsm) / rsm;
        }
      }
      // the candidates' own geometry, thread-parallel, ahead of the wavefront-per-candidate sums
      for (int c = tid; c < nc; c += blockDim.x) mh_geom_dst(R, rx, ry, rz, x, S.cand[c], x, S.gdst[c]);
      if (tid == blockDim.x - 1) mh_geom_src(R, rx, ry, rz, x, x, S.gsrc);
      __syncthreads();
      for (int c = wv; c < nc; c += nwv) {
        mh_features_rel(R, rx, ry, rz, qb, S.loc, nloc, x, S.cand[c], qk, -1, -1, nc, S.feat[c], x, VX, VY, VZ,
                        src_lds ? S.rs : nullptr, src_lds ? S.es : nullptr, S.gsrc, S.gdst[c]);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const float s = mlp_score_wave(S.w, S.feat[c]);
        if (lane == 0) S.q[c] = (double)s;
      }
      __syncthreads();
      if (wv == 0) mh_proposal_probs(S.q, nc, R.pfloor, &S.zq);
    } else if (tid == 0) {
      for (int c = 0; c < nc; ++c) S.q[c] = 1.0 / nc;
    }
    __syncthreads();
    if (tid == 0) {
      double tot = 0.0;
      for (int c = 0; c < nc; ++c) tot += S.q[c];
      const double t = (double)S.u[1] * tot;
      double cum = 0.0;
      int pick = nc - 1;
      for (int c = 0; c < nc; ++c) {
        cum += S.q[c];
        if (cum >= t) {
          pick = c;
          break;
        }
      }
      S.pick = pick;
      S.y = S.cand[pick];
      S.qf = S.q[pick];
    }
    __syncthreads();
    const int y = S.y;
    // features of the proposed move (training record)
    if (fa || learned) {
      if (learned) {
        if (tid < 16) S.fsel[tid] = S.feat[S.pick][tid];
      } else if (wv == 0) {
        mh_features_rel(R, rx, ry, rz, qb, S.loc, nloc, x, y, qk, -1, -1, nc, S.fsel, x, VX, VY, VZ, nullptr, nullptr);
      }
    }
    // exact energy change over all charges: V at the start of the round plus
    // this round's earlier moves, one of the four terms of one move per thread
    double dv = 0.0;
    {
      const int m = tid & 63, term = tid >> 6;
      if (m < S.nmv && term < 4) {
        const int4 mv = S.mv[m];
        const int p = (term < 2) ? y : x, b = (term & 1) ? mv.y : mv.z;
        const double ph = mh_phi(R, rx, ry, rz, p, b);
        dv = (double)mv.w * ((term == 0 || term == 3) ? ph : -ph);
      }
    }
    const double vy = V[R.slot0 + y];
    dv = block_sum(dv, S.dsum);
    const double de = R.lambda_b * (double)qk * (vy - vx + dv) - R.lambda_b * mh_phi(R, rx, ry, rz, x, y);
    // reverse proposal from y in the state after the move: y's hop list within
    // x's neighbourhood, less y and the beads charged after the move (one wavefront)
    if (wv == 0) {
      const int g0 = hop_off[R.slot0 + y], ng = hop_off[R.slot0 + y + 1] - g0;
      int cnt = 0;
      for (int base = 0; base < ng; base += 64) {
        const int i = base + lane;
        const int b = i < ng ? hop_list[g0 + i] : -1;
        bool f = false;
        if (b >= 0 && b != y) {
          const int qbb = (b == x) ? 0 : (int)qb[b];
          f = qbb == 0 && mh_dist2(R, rx, ry, rz, x, b) < R.nbhd2;
        }
        const unsigned long long m = __ballot(f);
        const int p = cnt + __popcll(m & ((1ull << lane) - 1ull));
        if (f && p < MH_MAXC) S.cand[p] = b;
        cnt += __popcll(m);
      }
      if (lane == 0) S.ncand = cnt;
    }
    __syncthreads();
    const int nc2 = S.ncand;
    if (nc2 > MH_MAXC) {
      if (tid == 0) atomicOr(err, 8);
      halted = true;
      break;
    }
    if (learned) {
      if (src_lds) {
        const int sy = R.slot0 + y, sx = R.slot0 + x;
        double vy3[3];
        ref_nid(R, rx[sy], ry[sy], rz[sy], rx[sx], ry[sx], rz[sx], vy3[0], vy3[1], vy3[2]);
        for (int i = tid; i < nloc; i += blockDim.x) {
          if (S.loc[i] == y) continue;
          const double rsm = mh_rel_dist(vy3[0], vy3[1], vy3[2], S.vx[i], S.vy[i], S.vz[i]);
          S.rs[i] = rsm;
          S.es[i] = exp(-R.kappa_d * rsm) / rsm;
        }
      }
      for (int c = tid; c < nc2; c += blockDim.x) mh_geom_dst(R, rx, ry, rz, y, S.cand[c], x, S.gdst[c]);
      if (tid == blockDim.x - 1) mh_geom_src(R, rx, ry, rz, y, x, S.gsrc);
      __syncthreads();
      for (int c = wv; c < nc2; c += nwv) {
        mh_features_rel(R, rx, ry, rz, qb, S.loc, nloc, y, S.cand[c], qk, x, y, nc2, S.feat[c], x, VX, VY, VZ,
                        src_lds ? S.rs : nullptr, src_lds ? S.es : nullptr, S.gsrc, S.gdst[c]);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const float s = mlp_score_wave(S.w, S.feat[c]);
        if (lane == 0) S.q[c] = (double)s;
      }
      __syncthreads();
      if (wv == 0) mh_proposal_probs(S.q, nc2, R.pfloor, &S.zq);
    }
    __syncthreads();
    if (tid == 0) {
      double qr = 0.0;
      for (int c = 0; c < nc2; ++c)
        if (S.cand[c] == x) qr = learned ? S.q[c] : 1.0 / nc2;
      const double lnrho = -de + log(qr) - log(S.qf);
      const bool adj = ((x / R.nb) == (y / R.nb)) && abs(y - x) == 1;
      const double ab = exp(-(adj ? R.barrier : R.barrier2));
      const double sr = exp(0.5 * lnrho);
      const double m1 = fmax(ab * sr, ab / sr);
      const double A = ab * sr / fmax(m1, 1.0);
      const int acc = (double)S.u[2] < A;
      S.accept = acc;
      old[R.q0 + k] = x;
      if (acc) {
        tg[k] = y;
        qg[x] = 0;
        qg[y] = (signed char)qk;
        if constexpr (LQ) {
          tv[k] = y;
          qb[x] = 0;
          qb[y] = (signed char)qk;
        }
        const int4 mv = make_int4(k, x, y, qk);
        S.mv[S.nmv] = mv;
        moves[(size_t)r * MH_ROUND + S.nmv] = mv;
        S.nmv = S.nmv + 1;
        if (qdisp) {
          double d[3];
          hop_vector(R, rx, ry, rz, R.slot0 + y, R.slot0 + x, d);
          for (int q = 0; q < 3; ++q) qdisp[3 * (R.q0 + k) + q] += d[q];
        }
      }
      if (lg) {
        lg[0] = k;
        lg[1] = x;
        lg[2] = y;
        lg[3] = adj ? 0 : 2;
        lg[4] = acc;
        lg[5] = 3;
      }
    }
    if (fa && tid < 16) fa[tid] = S.fsel[tid];
    done_att = a + 1;
    __syncthreads();
  }
  if (tid == 0) {
    nmoves[r] = S.nmv;
    done[r] = done_att;
    if (halted) stop[r] = 1;
    if (consumed) consumed[r] = 3 * done_att;
  }
}

void launch_mh_lists(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_n, const double* rx,
                     const double* ry, const double* rz, int* cnt, const int* off, int* list, int* hcnt,
                     const int* hoff, int* hlist, bool fill) {
  if (nrep <= 0 || max_n <= 0) return;
  const dim3 g((max_n + MC_THREADS / 64 - 1) / (MC_THREADS / 64), nrep);
  if (fill)
    k_mh_fill<<<g, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, off, list, hoff, hlist);
  else
    k_mh_count<<<g, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, cnt, hcnt);
}

void launch_mh_sweep(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_n, int max_nq,
                     const double* rx, const double* ry, const double* rz, int* track, int* old, const int* sign,
                     signed char* occ, long long occ_stride, const float* uniforms, int nu, const uint64_t* seeds,
                     uint64_t sweep_id, const float* weights, int* consumed, int* log, float* feats,
                     const int* log_off, int* err, double* qdisp, const int* nb_off, const int* nb_list,
                     const int* hop_off, const int* hop_list, double* V, void* moves, int* nmoves, int* done,
                     int* stop) {
  if (nrep <= 0) return;
  k_mh_begin<<<nrep, MC_THREADS, 0, st>>>(reps, rep_begin, track, sign, occ, occ_stride, done, stop, consumed);
  if (max_nq <= 0 || max_n <= 0) return;
  const dim3 g((max_n + MC_THREADS / 64 - 1) / (MC_THREADS / 64), nrep);
  k_mh_potential<<<g, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, track, sign, V);
  const bool lq = max_n <= MH_LDS_N && max_nq <= MH_LDS_Q;
  for (int a0 = 0; a0 < max_nq; a0 += MH_ROUND) {
    if (lq && weights)
      k_mh_round<true, MH_THREADS_LEARNED><<<nrep, MH_THREADS_LEARNED, 0, st>>>(
          reps, rep_begin, rx, ry, rz, track, old, sign, occ, occ_stride, uniforms, nu, seeds, sweep_id, weights,
          consumed, log, feats, log_off, err, qdisp, nb_off, nb_list, hop_off, hop_list, V, (int4*)moves, nmoves, done,
          stop, a0);
    else if (lq)
      k_mh_round<true, MH_THREADS_UNIFORM><<<nrep, MH_THREADS_UNIFORM, 0, st>>>(
          reps, rep_begin, rx, ry, rz, track, old, sign, occ, occ_stride, uniforms, nu, seeds, sweep_id, weights,
          consumed, log, feats, log_off, err, qdisp, nb_off, nb_list, hop_off, hop_list, V, (int4*)moves, nmoves, done,
          stop, a0);
    else if (weights)
      k_mh_round<false, MH_THREADS_LEARNED><<<nrep, MH_THREADS_LEARNED, 0, st>>>(
          reps, rep_begin, rx, ry, rz, track, old, sign, occ, occ_stride, uniforms, nu, seeds, sweep_id, weights,
          consumed, log, feats, log_off, err, qdisp, nb_off, nb_list, hop_off, hop_list, V, (int4*)moves, nmoves, done,
          stop, a0);
    else
      k_mh_round<false, MH_THREADS_UNIFORM><<<nrep, MH_THREADS_UNIFORM, 0, st>>>(
          reps, rep_begin, rx, ry, rz, track, old, sign, occ, occ_stride, uniforms, nu, seeds, sweep_id, weights,
          consumed, log, feats, log_off, err, qdisp, nb_off, nb_list, hop_off, hop_list, V, (int4*)moves, nmoves, done,
          stop, a0);
    if (a0 + MH_ROUND < max_nq)
      k_mh_fold<<<g, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, (const int4*)moves, nmoves, V);
  }
}
