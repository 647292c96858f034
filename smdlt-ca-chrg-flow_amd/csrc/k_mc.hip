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
// 2 hop + feature radius of x (all any forward or reverse feature can see) are
// gathered once per attempt in ascending order; features are summed by one
// wavefront per candidate.  dE is exact FP64 over all charges.
#define MH_MAXL 4096
#define MH_MAXC 128
#define MLP_NP (16 * 64 + 64 + 64 * 64 + 64 + 64 + 1)

struct MhShared {
  int loc[MH_MAXL];
  int cand[MH_MAXC];
  double q[MH_MAXC];
  float feat[MH_MAXC][16];
  float fsel[16];
  float w[MLP_NP];
  double dsum[MC_THREADS / 64];
  int wcnt[MC_THREADS / 64];
  int nloc, ncand, k, x, qk, y, pick, accept, bad;
  float u[3];
  double qf;
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

// 16 features of moving the charge at src (sign qk) to dst, summed by one
// wavefront over the gathered neighbourhood; clr / set (or -1) give the state
// after a move.  Lane 0 writes out[16].
__device__ void mh_features_wave(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                 const signed char* qb, const int* loc, int nloc, int src, int dst, int qk, int clr,
                                 int set, int ncand, float* out) {
  const int lane = threadIdx.x & 63;
  double e = 0, n4d = 0, n4s = 0, s4d = 0, s4s = 0, b4d = 0, b4s = 0;
  const double R4 = R.feat_r, kap = R.kappa_d;
  for (int i = lane; i < nloc; i += 64) {
    const int m = loc[i];
    if (m == src || m == dst) continue;
    const double rdm = mh_dist(R, rx, ry, rz, dst, m), rsm = mh_dist(R, rx, ry, rz, src, m);
    const int qm = (m == clr) ? 0 : (m == set) ? qk : (int)qb[m];
    if (qm != 0 && (rdm < R4 || rsm < R4))
      e += R.lambda_b * (double)(qk * qm) * (exp(-kap * rdm) / rdm - exp(-kap * rsm) / rsm);
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
    out[0] = (float)mh_dist(R, rx, ry, rz, src, dst);
    out[1] = (same && gap == 1) ? 1.f : 0.f;
    out[2] = same ? (float)((double)(gap < 20 ? gap : 20) / 20.0) : 1.f;
    out[3] = (float)e;
    out[4] = (float)n4d;
    out[5] = (float)n4s;
    out[6] = (float)s4d;
    out[7] = (float)s4s;
    out[8] = (float)b4d;
    out[9] = (float)b4s;
    out[10] = (float)mh_cos_at(R, rx, ry, rz, dst);
    out[11] = (float)mh_cos_at(R, rx, ry, rz, src);
    out[12] = (dst % nb == 0 || dst % nb == nb - 1) ? 1.f : 0.f;
    out[13] = (src % nb == 0 || src % nb == nb - 1) ? 1.f : 0.f;
    out[14] = (float)qk;
    out[15] = (float)((double)ncand / 10.0);
  }
}

// MLP score of one feature vector by one wavefront: lane j owns hidden unit j.
__device__ float mlp_score_wave(const float* w, const float* f) {
  const int j = threadIdx.x & 63;
  const float *W1 = w, *b1 = W1 + 64 * 16, *W2 = b1 + 64, *b2 = W2 + 64 * 64, *W3 = b2 + 64, *b3 = W3 + 64;
  float a = b1[j];
#pragma unroll
  for (int i = 0; i < 16; ++i) a = fmaf(W1[j * 16 + i], f[i], a);
  const float h1 = fmaxf(a, 0.f);
  float a2 = b2[j];
  for (int i = 0; i < 64; ++i) a2 = fmaf(W2[j * 64 + i], __shfl(h1, i, 64), a2);
  const float h2 = fmaxf(a2, 0.f);
  float s = W3[j] * h2;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s + b3[0];
}

// proposal over the scored candidates (thread 0): the network's predicted
// acceptance sigmoid(score) normalised (a Barker-type locally balanced weight),
// mixed with the uniform proposal at weight eps
__device__ void mh_proposal_probs(double* q, int nc, double eps) {
  double z = 0.0;
  for (int c = 0; c < nc; ++c) {
    q[c] = 1.0 / (1.0 + exp(-q[c]));
    z += q[c];
  }
  for (int c = 0; c < nc; ++c) q[c] = (1.0 - eps) * q[c] / z + eps / nc;
}

__global__ __launch_bounds__(MC_THREADS) void k_mh_sweep(const RepDev* __restrict__ reps, int rep_begin,
                                                         const double* __restrict__ rx, const double* __restrict__ ry,
                                                         const double* __restrict__ rz, int* track, int* old,
                                                         const int* sign, signed char* occ, long long occ_stride,
                                                         const float* uniforms, int nu, const uint64_t* seeds,
                                                         uint64_t sweep_id, const float* weights, int* consumed,
                                                         int* mlog, float* feats, const int* log_off, int* err,
                                                         double* qdisp) {
  const int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  __shared__ MhShared S;
  const int N = R.n, nq = R.nq, tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int nwv = blockDim.x >> 6;
  const bool learned = weights != nullptr;
  signed char* qb = occ + (size_t)blockIdx.x * occ_stride;
  for (int i = tid; i < N; i += blockDim.x) qb[i] = 0;
  if (learned)
    for (int i = tid; i < MLP_NP; i += blockDim.x) S.w[i] = weights[i];
  __syncthreads();
  for (int k = tid; k < nq; k += blockDim.x) qb[track[R.q0 + k]] = (signed char)sign[R.q0 + k];
  __syncthreads();
  const float* ur = uniforms ? uniforms + (size_t)r * nu : nullptr;
  int done_att = 0;
  for (int a = 0; a < nq; ++a) {
    if (tid == 0) {
      S.bad = 0;
      for (int i = 0; i < 3; ++i) {
        if (ur) {
          if (3 * a + i < nu) {
            S.u[i] = ur[3 * a + i];
          } else {
            S.bad = 1;
            S.u[i] = 0.f;
          }
        } else {
          S.u[i] = dev_uniform(seeds[r], sweep_id, (uint32_t)(3 * a + i));
        }
      }
      int k = (int)(S.u[0] * (float)nq);
      S.k = k > nq - 1 ? nq - 1 : k;
      S.x = track[R.q0 + S.k];
      S.qk = sign[R.q0 + S.k];
      S.nloc = 0;
      S.ncand = 0;
    }
    __syncthreads();
    if (S.bad) {
      if (tid == 0) atomicOr(err, 2);
      break;
    }
    const int k = S.k, x = S.x, qk = S.qk;
    int* lg = mlog ? mlog + 6 * (size_t)(log_off[r] + a) : nullptr;
    float* fa = feats ? feats + 16 * (size_t)(log_off[r] + a) : nullptr;
    // neighbourhood of x (x included), ascending
    for (int base = 0; base < N; base += blockDim.x) {
      const int b = base + tid;
      const bool f = b < N && mh_dist2(R, rx, ry, rz, x, b) < R.nbhd2;
      block_append(f, b, S.loc, &S.nloc, MH_MAXL, S.wcnt);
    }
    const int nloc = S.nloc;
    if (nloc > MH_MAXL) {
      if (tid == 0) atomicOr(err, 8);
      break;
    }
    // candidates of x
    for (int base = 0; base < nloc; base += blockDim.x) {
      const int i = base + tid;
      const int b = i < nloc ? S.loc[i] : -1;
      const bool f = b >= 0 && b != x && qb[b] == 0 && mh_dist2(R, rx, ry, rz, x, b) < R.hop2;
      block_append(f, b, S.cand, &S.ncand, MH_MAXC, S.wcnt);
    }
    const int nc = S.ncand;
    if (nc > MH_MAXC) {
      if (tid == 0) atomicOr(err, 8);
      break;
    }
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
      for (int c = wv; c < nc; c += nwv) {
        mh_features_wave(R, rx, ry, rz, qb, S.loc, nloc, x, S.cand[c], qk, -1, -1, nc, S.feat[c]);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const float s = mlp_score_wave(S.w, S.feat[c]);
        if (lane == 0) S.q[c] = (double)s;
      }
      __syncthreads();
      if (tid == 0) mh_proposal_probs(S.q, nc, R.pfloor);
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
        mh_features_wave(R, rx, ry, rz, qb, S.loc, nloc, x, y, qk, -1, -1, nc, S.fsel);
      }
    }
    // exact energy change over all charges
    double de = 0.0;
    for (int m = tid; m < nq; m += blockDim.x) {
      if (m == k) continue;
      const int bm = track[R.q0 + m], qm = sign[R.q0 + m];
      const double rym = mh_dist(R, rx, ry, rz, y, bm), rxm = mh_dist(R, rx, ry, rz, x, bm);
      de += R.lambda_b * (double)(qk * qm) * (exp(-R.kappa_d * rym) / rym - exp(-R.kappa_d * rxm) / rxm);
    }
    de = block_sum(de, S.dsum);
    // reverse proposal from y in the state after the move
    if (tid == 0) S.ncand = 0;
    __syncthreads();
    for (int base = 0; base < nloc; base += blockDim.x) {
      const int i = base + tid;
      const int b = i < nloc ? S.loc[i] : -1;
      bool f = false;
      if (b >= 0 && b != y) {
        const int qbb = (b == x) ? 0 : (int)qb[b];
        f = qbb == 0 && mh_dist2(R, rx, ry, rz, y, b) < R.hop2;
      }
      block_append(f, b, S.cand, &S.ncand, MH_MAXC, S.wcnt);
    }
    const int nc2 = S.ncand;
    if (nc2 > MH_MAXC) {
      if (tid == 0) atomicOr(err, 8);
      break;
    }
    if (learned) {
      for (int c = wv; c < nc2; c += nwv) {
        mh_features_wave(R, rx, ry, rz, qb, S.loc, nloc, y, S.cand[c], qk, x, y, nc2, S.feat[c]);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const float s = mlp_score_wave(S.w, S.feat[c]);
        if (lane == 0) S.q[c] = (double)s;
      }
      __syncthreads();
      if (tid == 0) mh_proposal_probs(S.q, nc2, R.pfloor);
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
        track[R.q0 + k] = y;
        qb[x] = 0;
        qb[y] = (signed char)qk;
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
  if (tid == 0 && consumed) consumed[r] = 3 * done_att;
}

void launch_mh_sweep(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* rx, const double* ry,
                     const double* rz, int* track, int* old, const int* sign, signed char* occ, long long occ_stride,
                     const float* uniforms, int nu, const uint64_t* seeds, uint64_t sweep_id, const float* weights,
                     int* consumed, int* log, float* feats, const int* log_off, int* err, double* qdisp) {
  if (nrep <= 0) return;
  k_mh_sweep<<<nrep, MC_THREADS, 0, st>>>(reps, rep_begin, rx, ry, rz, track, old, sign, occ, occ_stride, uniforms, nu,
                                          seeds, sweep_id, weights, consumed, log, feats, log_off, err, qdisp);
}
