// Neighbour lists from a cell grid in the Lagrange-reduced lattice basis.
// One build produces, per bead, (a) the real-space mobility list: every bead
// within rc+skin under the true minimum image, and (b) the reference's Verlet
// list: every bead with |getNID| < rv, bit-exact with verletlist
// (S/main26.c:682-707) because the candidate test is the reference's FP64
// getNID.  Cells are at least max(rc+skin, rv+1) wide, so both candidate sets
// lie in the 27 surrounding cells (the +1 covers the shear delx shift).
// Lists are column-major ([k*nslots + i]) and sorted by bead index.
#include "common.h"

__device__ __forceinline__ void cell_coords(const RepDev& R, double x, double y, double z, int c[3]) {
  double f[3];
  f[0] = R.rd[0] * x + R.rd[1] * y;
  f[1] = R.rd[2] * x + R.rd[3] * y;
  f[2] = z * R.invL;
  for (int d = 0; d < 3; ++d) {
    double g = f[d] - floor(f[d]);
    int k = (int)(g * R.ncell[d]);
    if (k >= R.ncell[d]) k = R.ncell[d] - 1;
    if (k < 0) k = 0;
    c[d] = k;
  }
}

// Cells are only read by the list builders, so a replica whose lists are kept
// this step (rep_flags 0) is not binned: its cells stay empty.
__global__ void k_cell_count(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                             int nslots, const double* rx, const double* ry, const double* rz, int* cell_of,
                             int* cell_cnt, const int* __restrict__ rep_flags) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || !(rep_flags[r] & 3)) return;
  if (s - R.slot0 >= R.n) {  // padding slot: never binned
    cell_of[s] = -1;
    return;
  }
  int c[3];
  cell_coords(R, rx[s], ry[s], rz[s], c);
  int id = R.cell0 + (c[2] * R.ncell[1] + c[1]) * R.ncell[0] + c[0];
  cell_of[s] = id;
  atomicAdd(&cell_cnt[id], 1);
}

void launch_cell_count(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                       const double* rx, const double* ry, const double* rz, int* cell_of, int* cell_cnt,
                       const int* rep_flags) {
  if (nslots <= 0) return;
  k_cell_count<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, rx, ry, rz, cell_of,
                                                     cell_cnt, rep_flags);
}

// Exclusive scan of n ints by one workgroup (chunked per thread).
__global__ void k_scan(const int* in, int* out, int n) {
  __shared__ int part[1024];
  int t = threadIdx.x, T = blockDim.x;
  int chunk = (n + T - 1) / T;
  int b = t * chunk, e = min(n, b + chunk);
  int s = 0;
  for (int i = b; i < e; ++i) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < T; o <<= 1) {
    int v = (t >= o) ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = b; i < e; ++i) {
    int v = in[i];
    out[i] = run;
    run += v;
  }
}

void launch_scan(hipStream_t st, const int* in, int* out, int n) {
  if (n <= 0) return;
  k_scan<<<1, 1024, 0, st>>>(in, out, n);
}

__global__ void k_cell_fill(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                            int nslots, const int* cell_of, const int* cell_start, int* cell_fill, int* cell_beads,
                            const int* __restrict__ rep_flags) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0 || !reps[r].active || s - reps[r].slot0 >= reps[r].n || !(rep_flags[r] & 3)) return;
  int id = cell_of[s];
  if (id < 0) return;
  int k = atomicAdd(&cell_fill[id], 1);
  cell_beads[cell_start[id] + k] = s;
}

void launch_cell_fill(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                      const int* cell_of, const int* cell_start, int* cell_fill, int* cell_beads,
                      const int* rep_flags) {
  if (nslots <= 0) return;
  k_cell_fill<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, cell_of, cell_start,
                                                    cell_fill, cell_beads, rep_flags);
}

// Writes each cell's members in ascending slot order: a member's place is the number of
// members with a smaller slot.  k_cell_fill's atomics leave the order in which they landed,
// and list rows, with the FP64 sums that follow them, inherit the cell order, so this makes
// every result independent of that order.  One workgroup per cell; members are staged
// through LDS in chunks, so any cell size works.
constexpr int CELL_SORT_CHUNK = 2048;

__global__ __launch_bounds__(256) void k_cell_sort(const int* __restrict__ cell_start,
                                                   const int* __restrict__ cell_cnt, int cell_begin,
                                                   const int* __restrict__ unsorted, int* __restrict__ cell_beads) {
  __shared__ int v[CELL_SORT_CHUNK];
  const int id = cell_begin + blockIdx.x;
  const int m = cell_cnt[id], base = cell_start[id];
  const int* src = unsorted + base;
  if (m == 1 && threadIdx.x == 0) cell_beads[base] = src[0];
  if (m <= 1) return;
  for (int k0 = 0; k0 < m; k0 += blockDim.x) {
    const int k = k0 + threadIdx.x;
    const int x = k < m ? src[k] : 0;
    int rank = 0;
    for (int q0 = 0; q0 < m; q0 += CELL_SORT_CHUNK) {
      const int qe = min(CELL_SORT_CHUNK, m - q0);
      __syncthreads();
      for (int q = threadIdx.x; q < qe; q += blockDim.x) v[q] = src[q0 + q];
      __syncthreads();
      for (int q = 0; q < qe; ++q) rank += v[q] < x;
    }
    if (k < m) cell_beads[base + rank] = x;
  }
}

void launch_cell_sort(hipStream_t st, const int* cell_start, const int* cell_cnt, int cell_begin, int ncells,
                      const int* unsorted, int* cell_beads) {
  if (ncells <= 0) return;
  k_cell_sort<<<ncells, 256, 0, st>>>(cell_start, cell_cnt, cell_begin, unsorted, cell_beads);
}

__device__ __forceinline__ void sorted_insert(int* list, int stride, int i, int& cnt, int cap, int j, int* overflow) {
  if (cnt >= cap) {
    atomicOr(overflow, 1);
    return;
  }
  int pos = cnt;
  while (pos > 0) {
    int prev = list[(size_t)(pos - 1) * stride + i];
    if (prev <= j) break;
    list[(size_t)pos * stride + i] = prev;
    --pos;
  }
  list[(size_t)pos * stride + i] = j;
  ++cnt;
}

__device__ __forceinline__ void neighbour_cell_range(const RepDev& R, int lo[3], int hi[3]) {
  for (int d = 0; d < 3; ++d) {
    int n = R.ncell[d];
    lo[d] = (n >= 3) ? -1 : 0;  // with fewer than 3 cells visit each distinct cell once
    hi[d] = (n >= 3) ? 1 : n - 1;
  }
}

// Mobility list (rep_flags bit1): one wavefront per bead, lanes test the
// candidates of each neighbouring cell and append hits in wave order (ballot).
__global__ __launch_bounds__(256) void k_build_rpy(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep,
                                                   int slot_begin, int nslots, const double* rx, const double* ry,
                                                   const double* rz, const int* cell_start, const int* cell_cnt,
                                                   const int* cell_beads, const int* rep_flags, int nst, int cap,
                                                   int* nrpy, int* lrpy, int* overflow) {
  const int lane = threadIdx.x & 63;
  int i = slot_begin + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= slot_begin + nslots) return;
  int r = slot_rep[i];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || !(rep_flags[r] & 2)) return;
  if (i - R.slot0 >= R.n) {
    if (lane == 0) nrpy[i] = 0;
    return;
  }
  const double xi = rx[i], yi = ry[i], zi = rz[i];
  int c[3], lo[3], hi[3];
  cell_coords(R, xi, yi, zi, c);
  neighbour_cell_range(R, lo, hi);
  int cr = 0;
  for (int o3 = lo[2]; o3 <= hi[2]; ++o3)
    for (int o2 = lo[1]; o2 <= hi[1]; ++o2)
      for (int o1 = lo[0]; o1 <= hi[0]; ++o1) {
        int a = (c[0] + o1 + R.ncell[0]) % R.ncell[0];
        int b = (c[1] + o2 + R.ncell[1]) % R.ncell[1];
        int e = (c[2] + o3 + R.ncell[2]) % R.ncell[2];
        int id = R.cell0 + (e * R.ncell[1] + b) * R.ncell[0] + a;
        int beg = cell_start[id], cnt = cell_cnt[id];
        for (int k0 = 0; k0 < cnt; k0 += 64) {
          int k = k0 + lane;
          bool hit = false;
          int j = -1;
          if (k < cnt) {
            j = cell_beads[beg + k];
            if (j != i) {
              double dx = rx[j] - xi, dy = ry[j] - yi, dz = rz[j] - zi;
              min_image_reduced(R, dx, dy, dz);
              hit = dx * dx + dy * dy + dz * dz < R.rlist2;
            }
          }
          unsigned long long m = __ballot(hit);
          int pos = cr + __popcll(m & ((1ull << lane) - 1ull));
          if (hit) {
            if (pos < cap)
              lrpy[(size_t)i * cap + pos] = j;
            else
              atomicOr(overflow, 1);
          }
          cr += __popcll(m);
        }
      }
  if (lane == 0) nrpy[i] = min(cr, cap);
}

// ascending bitonic sort of one value per lane across a wavefront
__device__ __forceinline__ int wave_sort(int v, int lane) {
  for (int k = 2; k <= 64; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      int o = __shfl_xor(v, j, 64);
      bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
      v = keep_min ? min(v, o) : max(v, o);
    }
  return v;
}

// The reference's Verlet list (rep_flags bit0): all j with |getNID(i,j)| < rv,
// ascending.  One wavefront per bead: lanes test candidates (from the mobility
// list when its cutoff covers rv + 1 -- the delx shift changes a pair distance
// by less than 1 -- otherwise from the cell grid), hits are compacted in LDS
// and sorted (bitonic across the wavefront, or serially past 64 hits).
#define LJ_BUF 256
__global__ __launch_bounds__(256) void k_build_lj(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep,
                                                  int slot_begin, int nslots, const double* rx, const double* ry,
                                                  const double* rz, const int* cell_start, const int* cell_cnt,
                                                  const int* cell_beads, const int* rep_flags, int nst,
                                                  const int* nrpy, const int* lrpy, int cap_rpy, int cap_lj,
                                                  int* nlj, int* llj, int* overflow) {
  __shared__ int buf[4][LJ_BUF];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int i = slot_begin + blockIdx.x * 4 + wv;
  if (i >= slot_begin + nslots) return;
  int r = slot_rep[i];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || !(rep_flags[r] & 1)) return;
  if (i - R.slot0 >= R.n) {
    if (lane == 0) nlj[i] = 0;
    return;
  }
  const double xi = rx[i], yi = ry[i], zi = rz[i];
  int cl = 0;
  auto test = [&](int j, bool valid) {
    bool hit = false;
    if (valid && j != i) {
      double ox, oy, oz;
      ref_nid(R, xi, yi, zi, rx[j], ry[j], rz[j], ox, oy, oz);
      double rr;
      {
#pragma clang fp contract(off)
        rr = sqrt(ox * ox + oy * oy + oz * oz);
      }
      hit = rr < R.rv;
    }
    unsigned long long m = __ballot(hit);
    int pos = cl + __popcll(m & ((1ull << lane) - 1ull));
    if (hit && pos < LJ_BUF) buf[wv][pos] = j;
    cl += __popcll(m);
  };
  if (R.rc2 >= (R.rv + 1.0) * (R.rv + 1.0)) {
    int m = nrpy[i];
    for (int k0 = 0; k0 < m; k0 += 64) {
      int k = k0 + lane;
      test(k < m ? lrpy[(size_t)i * cap_rpy + k] : -1, k < m);
    }
  } else {
    int c[3], lo[3], hi[3];
    cell_coords(R, xi, yi, zi, c);
    neighbour_cell_range(R, lo, hi);
    for (int o3 = lo[2]; o3 <= hi[2]; ++o3)
      for (int o2 = lo[1]; o2 <= hi[1]; ++o2)
        for (int o1 = lo[0]; o1 <= hi[0]; ++o1) {
          int a = (c[0] + o1 + R.ncell[0]) % R.ncell[0];
          int b = (c[1] + o2 + R.ncell[1]) % R.ncell[1];
          int e = (c[2] + o3 + R.ncell[2]) % R.ncell[2];
          int id = R.cell0 + (e * R.ncell[1] + b) * R.ncell[0] + a;
          int beg = cell_start[id], cnt = cell_cnt[id];
          for (int k0 = 0; k0 < cnt; k0 += 64) {
            int k = k0 + lane;
            test(k < cnt ? cell_beads[beg + k] : -1, k < cnt);
          }
        }
  }
  if (cl > LJ_BUF || cl > cap_lj) {
    if (lane == 0) atomicOr(overflow, 1);
    cl = min(cl, min(LJ_BUF, cap_lj));
  }
  if (cl <= 64) {
    int v = (lane < cl) ? buf[wv][lane] : INT_MAX;
    v = wave_sort(v, lane);
    if (lane < cl) llj[(size_t)lane * nst + i] = v;
  } else if (lane == 0) {
    int* q = buf[wv];
    for (int a = 1; a < cl; ++a) {
      int x = q[a], b = a - 1;
      while (b >= 0 && q[b] > x) {
        q[b + 1] = q[b];
        --b;
      }
      q[b + 1] = x;
    }
    for (int a = 0; a < cl; ++a) llj[(size_t)a * nst + i] = q[a];
  }
  if (lane == 0) nlj[i] = cl;
}

void launch_build_lists(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                        const double* rx, const double* ry, const double* rz, const int* cell_start,
                        const int* cell_cnt, const int* cell_beads, const int* rep_flags, int nslots_total,
                        int cap_rpy, int* nrpy, int* lrpy, int cap_lj, int* nlj, int* llj, int* overflow) {
  if (nslots <= 0) return;
  k_build_rpy<<<(nslots + 3) / 4, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, rx, ry, rz, cell_start, cell_cnt,
                                                cell_beads, rep_flags, nslots_total, cap_rpy, nrpy, lrpy, overflow);
  k_build_lj<<<(nslots + 3) / 4, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, rx, ry, rz, cell_start, cell_cnt,
                                               cell_beads, rep_flags, nslots_total, nrpy, lrpy, cap_rpy, cap_lj, nlj,
                                               llj, overflow);
}

// Cluster-pair list for the mixed-precision real-space pass: cluster J is listed
// for cluster I when some bead of J is in the mobility list of some bead of I
// (pairs within rc + skin at build time, so every pair within rc until the next
// rebuild is covered), plus I itself.  One wavefront per cluster marks the
// clusters in an LDS bitmap, then emits them in ascending order, each entry
// packed as (first bead slot << 4) | bead count.
__global__ __launch_bounds__(64) void k_build_clusters(const RepDev* __restrict__ reps,
                                                        const int* __restrict__ slot_rep, int cl_begin, int ncl,
                                                        const int* cl_first, const int* cl_cnt, const int* slot_cl,
                                                        const int* rep_flags, const int* nrpy, const int* lrpy,
                                                        int cap_rpy, int capc, int* nclp, int* lcl, int* overflow) {
  extern __shared__ unsigned bm[];
  const int I = cl_begin + blockIdx.x;
  if (I >= cl_begin + ncl) return;
  const int first = cl_first[I], cnt = cl_cnt[I];
  const int r = slot_rep[first];
  const RepDev& R = reps[r];
  if (!R.active || !(rep_flags[r] & 2)) return;
  const int lane = threadIdx.x;
  const int nw = (R.ncl + 31) >> 5;
  for (int w = lane; w < nw; w += 64) bm[w] = 0u;
  __syncthreads();
  if (lane == 0) atomicOr(&bm[(I - R.cl0) >> 5], 1u << ((I - R.cl0) & 31));
  for (int ii = 0; ii < cnt; ++ii) {
    const int i = first + ii, m = nrpy[i];
    for (int k = lane; k < m; k += 64) {
      const int cj = slot_cl[lrpy[(size_t)i * cap_rpy + k]] - R.cl0;
      atomicOr(&bm[cj >> 5], 1u << (cj & 31));
    }
  }
  __syncthreads();
  int base = 0;
  for (int w0 = 0; w0 < nw; w0 += 64) {
    const int w = w0 + lane;
    unsigned bits = w < nw ? bm[w] : 0u;
    int c = __popc(bits), incl = c;
    for (int o = 1; o < 64; o <<= 1) {
      int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    int pos = base + incl - c;
    while (bits) {
      const int b = __ffs(bits) - 1;
      bits &= bits - 1u;
      if (pos < capc) {
        const int J = R.cl0 + w * 32 + b;
        lcl[(size_t)I * capc + pos] = (cl_first[J] << 4) | cl_cnt[J];  // first bead, count: one load per partner
      } else
        atomicOr(overflow, 1);
      ++pos;
    }
    base += __shfl(incl, 63, 64);
  }
  if (lane == 0) nclp[I] = min(base, capc);
}

void launch_build_clusters(hipStream_t st, const RepDev* reps, const int* slot_rep, int cl_begin, int ncl,
                           int max_words, const int* cl_first, const int* cl_cnt, const int* slot_cl,
                           const int* rep_flags, const int* nrpy, const int* lrpy, int cap_rpy, int capc, int* nclp,
                           int* lcl, int* overflow) {
  if (ncl <= 0) return;
  k_build_clusters<<<ncl, 64, 4 * (size_t)max_words, st>>>(reps, slot_rep, cl_begin, ncl, cl_first, cl_cnt, slot_cl,
                                                          rep_flags, nrpy, lrpy, cap_rpy, capc, nclp, lcl, overflow);
}

// Fill kernels (used instead of memset so that captured graphs replay correctly).
template <typename T> __global__ void k_fill(T* p, T v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}
static int fill_blocks(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}
__global__ void k_copy_f64(double* dst, const double* src, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
void launch_copy_f64(hipStream_t st, double* dst, const double* src, long long n) {
  if (n > 0) k_copy_f64<<<fill_blocks(n), 256, 0, st>>>(dst, src, n);
}
void launch_fill_i32(hipStream_t st, int* p, int v, long long n) {
  if (n > 0) k_fill<int><<<fill_blocks(n), 256, 0, st>>>(p, v, n);
}
void launch_fill_f64(hipStream_t st, double* p, double v, long long n) {
  if (n > 0) k_fill<double><<<fill_blocks(n), 256, 0, st>>>(p, v, n);
}
void launch_fill_f32(hipStream_t st, float* p, float v, long long n) {
  if (n > 0) k_fill<float><<<fill_blocks(n), 256, 0, st>>>(p, v, n);
}

// Rebuild decision per replica: bit0 when the reference's Verlet trigger fires
// (max |dxv| > (rv-rc)/2, checkVerlet S/main26.c:858-869); bit1 when the
// mobility list may have lost a pair: 2 max|dsp| plus the affine stretch of a
// pair separation since the last build reaches the skin, or a new lattice period.
__global__ void k_nl_check(const RepDev* __restrict__ reps, int rep_begin, int nrep, const double* out2, double* aff,
                           int force, int* rep_flags) {
  int r = rep_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rep_begin + nrep) return;
  const RepDev& R = reps[r];
  if (!R.active) {
    rep_flags[r] = 0;
    return;
  }
  int f = 0;
  if (sqrt(out2[2 * r]) > R.rnew || (force & 1)) f |= 1;
  double a = aff[r] + fabs(R.edot) * R.dt;
  double reach = sqrt(R.rlist2);
  if (2.0 * sqrt(out2[2 * r + 1]) + a * reach >= R.skin || R.reset || (force & 2)) {
    f |= 2;
    a = 0.0;
  }
  aff[r] = a;
  rep_flags[r] = f;
}

void launch_nl_check(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* out2, double* aff,
                     int force, int* rep_flags) {
  if (nrep > 0) k_nl_check<<<(nrep + 63) / 64, 64, 0, st>>>(reps, rep_begin, nrep, out2, aff, force, rep_flags);
}

// Zero the displacement trackers of the replicas whose lists were rebuilt
// (verletlist zeroes dxv, S/main26.c:686-688).
__global__ void k_reset_trackers(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                                 int nslots, const int* rep_flags, double* dxv, double* dyv, double* dzv,
                                 double* dsp) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0 || !reps[r].active) return;
  int f = rep_flags[r];
  if (f & 1) dxv[s] = dyv[s] = dzv[s] = 0.0;
  if (f & 2) dsp[3 * s] = dsp[3 * s + 1] = dsp[3 * s + 2] = 0.0;
}

void launch_reset_trackers(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                           const int* rep_flags, double* dxv, double* dyv, double* dzv, double* dsp) {
  if (nslots > 0)
    k_reset_trackers<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, rep_flags, dxv, dyv,
                                                           dzv, dsp);
}
