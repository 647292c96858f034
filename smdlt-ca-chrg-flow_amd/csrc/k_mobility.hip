// Ewald-summed RPY mobility product u = M v for every replica in one launch
// sequence, replacing the dense matrix of calcHI/updateHI (S/main26.c:1530-1806,
// 745-857) with a smooth-PME split:
//   u_i = selfconst v_i + sum_{j in list(i), r<rc} M_real(r_ij) v_j + interp_i(phi),
//   phi = IDFT[ G(m) |b(m)|^2 DFT[ spread(v) ] ],
// on a K1 x K2 x K3 grid along the (deformed) lattice vectors, cardinal
// B-splines of order p (Essmann et al. 1995), and the Beenakker reciprocal
// kernel G(k) = (1 - k^2/3)(1 + k^2/4a^2 + k^4/8a^4)(6 pi/k^2) e^{-k^2/4a^2}
// (I - k k^T/k^2)/V (S/main26.c:1622).  The FFTs are hand-written in-LDS
// Stockham radix-4/2 passes: x (two real rows packed per complex line),
// y, and z with the Green multiply and inverse z fused; then inverse y and x.
// T = float (mixed mode; per-bead sums compensated) or double.
#include "common.h"

template <typename T> struct Prec {
  typedef typename cplx_t<T>::type C;
};

// -------------------------------------------------------------------- FFT core
// One in-place Stockham stage of radix R (2, 3, 4, 5, 8 or 9) over nl lines of length
// K in LDS: every thread first loads its butterflies into registers, then
// (after a barrier) writes them to their autosorted positions.  tw holds
// exp(-2 pi i t / K), t < K, for this line length.
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(a.x) s) {
  C r;
  r.x = a.x * s;
  r.y = a.y * s;
  return r;
}
// multiply by -i (forward) or +i (inverse)
template <typename C> __device__ __forceinline__ C crot(C d, int dir) {
  C r;
  if (dir < 0) {
    r.x = d.y;
    r.y = -d.x;
  } else {
    r.x = -d.y;
    r.y = d.x;
  }
  return r;
}

template <typename C> __device__ __forceinline__ C make_c(decltype(C().x) x, decltype(C().x) y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}
// 3-point DFT: y0 = a + (b + c), y1,2 = a - (b + c)/2 -+ i sin(2 pi/3) (b - c) (forward)
template <typename C>
__device__ __forceinline__ void dft3(C x0, C x1, C x2, int dir, C& y0, C& y1, C& y2) {
  typedef decltype(C().x) T;
  const T s3 = (T)0.86602540378443864676;
  C sm = cadd(x1, x2), df = csub(x1, x2);
  C m = csub(x0, cscale(sm, (T)0.5));
  C rt = cscale(crot(df, dir), s3);
  y0 = cadd(x0, sm);
  y1 = cadd(m, rt);
  y2 = csub(m, rt);
}

template <typename C, int R, int MAXB>
__device__ __forceinline__ void fft_stage(C* sh, int K, int pitch, int nl, int Ns, int dir, const C* __restrict__ tw) {
  typedef decltype(C().x) T;
  const int TH = blockDim.x, t = threadIdx.x;
  const int KR = K / R;
  const int nb = nl * KR;
  const FDiv dKR(KR), dNs(Ns);
  C v[MAXB][R];
  int lq[MAXB], jq[MAXB];  // each butterfly's line and position, kept for the write
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    int b = t + q * TH;
    if (b < nb) {
      int l = dKR.div(b), j = b - l * KR;
      lq[q] = l;
      jq[q] = j;
      const C* src = sh + l * pitch + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[r * KR];
    }
  }
  __syncthreads();
  const int tstep = K / (Ns * R);
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    int b = t + q * TH;
    if (b < nb) {
      const int l = lq[q], j = jq[q];
      int jm = j - dNs.div(j) * Ns;
      if (Ns > 1) {  // the first stage's twiddles are all 1
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C w = tw[jm * r * tstep];
          if (dir > 0) w = cconj(w);
          v[q][r] = cmul(v[q][r], w);
        }
      }
      C* dst = sh + l * pitch + (j - jm) * R + jm;
      if (R == 2) {
        dst[0] = cadd(v[q][0], v[q][1]);
        dst[Ns] = csub(v[q][0], v[q][1]);
      } else if (R == 4) {
        C a = cadd(v[q][0], v[q][R > 2 ? 2 : 0]), bb = csub(v[q][0], v[q][R > 2 ? 2 : 0]);
        C c = cadd(v[q][1], v[q][R > 3 ? 3 : 1]), d = csub(v[q][1], v[q][R > 3 ? 3 : 1]);
        C id = crot(d, dir);  // -i*d for forward, +i*d for inverse
        dst[0] = cadd(a, c);
        dst[Ns] = cadd(bb, id);
        dst[2 * Ns] = csub(a, c);
        dst[3 * Ns] = csub(bb, id);
      } else if (R == 3) {
        // y0 = a + (b + c); y1,2 = a - (b + c)/2 -+ i sin(2pi/3) (b - c) (forward)
        const T s3 = (T)0.86602540378443864676;
        C x0 = v[q][0], x1 = v[q][1], x2 = v[q][R > 2 ? 2 : 0];
        C sm = cadd(x1, x2), df = csub(x1, x2);
        C m = csub(x0, cscale(sm, (T)0.5));
        C rt = cscale(crot(df, dir), s3);
        dst[0] = cadd(x0, sm);
        dst[Ns] = cadd(m, rt);
        dst[2 * Ns] = csub(m, rt);
      } else if (R == 8) {
        // radix-2 decimation in frequency of the 8-point DFT: even outputs from
        // b_n = x_n + x_{n+4}, odd outputs from d_n = (x_n - x_{n+4}) W^n, W = e^{-+2 pi i/8}
        const T h = (T)0.70710678118654752440;
        C x0 = v[q][0], x1 = v[q][1], x2 = v[q][R > 2 ? 2 : 0], x3 = v[q][R > 3 ? 3 : 0];
        C x4 = v[q][R > 4 ? 4 : 0], x5 = v[q][R > 5 ? 5 : 0], x6 = v[q][R > 6 ? 6 : 0], x7 = v[q][R > 7 ? 7 : 0];
        C b0 = cadd(x0, x4), b1 = cadd(x1, x5), b2 = cadd(x2, x6), b3 = cadd(x3, x7);
        C d0 = csub(x0, x4), d1 = csub(x1, x5), d2 = crot(csub(x2, x6), dir), d3 = csub(x3, x7);
        C t1, t3;  // d1 W, d3 W^3 with W = (1 -+ i)/sqrt2
        if (dir < 0) {
          t1.x = h * (d1.x + d1.y);
          t1.y = h * (d1.y - d1.x);
          t3.x = h * (d3.y - d3.x);
          t3.y = -h * (d3.x + d3.y);
        } else {
          t1.x = h * (d1.x - d1.y);
          t1.y = h * (d1.x + d1.y);
          t3.x = -h * (d3.x + d3.y);
          t3.y = h * (d3.x - d3.y);
        }
        C e0 = cadd(b0, b2), e1 = cadd(b1, b3), f0 = csub(b0, b2), f1 = crot(csub(b1, b3), dir);
        C g0 = cadd(d0, d2), g1 = cadd(t1, t3), k0 = csub(d0, d2), k1 = crot(csub(t1, t3), dir);
        dst[0] = cadd(e0, e1);
        dst[4 * Ns] = csub(e0, e1);
        dst[2 * Ns] = cadd(f0, f1);
        dst[6 * Ns] = csub(f0, f1);
        dst[Ns] = cadd(g0, g1);
        dst[5 * Ns] = csub(g0, g1);
        dst[3 * Ns] = cadd(k0, k1);
        dst[7 * Ns] = csub(k0, k1);
      } else if constexpr (R == 9) {
        // 3 x 3: DFT3 over n1 of x[3 n1 + n2], twiddles W^(n2 k1), DFT3 over n2;
        // output k1 + 3 k2, W = e^{-+2 pi i/9}
        C A[3][3];
#pragma unroll
        for (int n2 = 0; n2 < 3; ++n2) dft3<C>(v[q][n2], v[q][3 + n2], v[q][6 + n2], dir, A[n2][0], A[n2][1], A[n2][2]);
        const T c1 = (T)0.76604444311897803520, s1 = (T)0.64278760968653932632;  // 2 pi/9
        const T c2 = (T)0.17364817766693034885, s2 = (T)0.98480775301220805936;  // 4 pi/9
        const T c4 = (T)-0.93969262078590838405, s4 = (T)0.34202014332566873304; // 8 pi/9
        const T sg = dir < 0 ? (T)-1 : (T)1;
        A[1][1] = cmul(A[1][1], make_c<C>(c1, sg * s1));
        A[1][2] = cmul(A[1][2], make_c<C>(c2, sg * s2));
        A[2][1] = cmul(A[2][1], make_c<C>(c2, sg * s2));
        A[2][2] = cmul(A[2][2], make_c<C>(c4, sg * s4));
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
          C y0, y1, y2;
          dft3<C>(A[0][k1], A[1][k1], A[2][k1], dir, y0, y1, y2);
          dst[k1 * Ns] = y0;
          dst[(k1 + 3) * Ns] = y1;
          dst[(k1 + 6) * Ns] = y2;
        }
      } else if (R == 5) {
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        C x0 = v[q][0], x1 = v[q][1], x2 = v[q][R > 2 ? 2 : 0], x3 = v[q][R > 3 ? 3 : 0], x4 = v[q][R > 4 ? 4 : 0];
        C a1 = cadd(x1, x4), b1 = csub(x1, x4), a2 = cadd(x2, x3), b2 = csub(x2, x3);
        C m1 = cadd(x0, cadd(cscale(a1, c1), cscale(a2, c2)));
        C m2 = cadd(x0, cadd(cscale(a1, c2), cscale(a2, c1)));
        // forward: y_k = m_k - i (s-combination); crot gives -i for forward
        C r1 = crot(cadd(cscale(b1, s1), cscale(b2, s2)), dir);
        C r2 = crot(csub(cscale(b1, s2), cscale(b2, s1)), dir);
        dst[0] = cadd(x0, cadd(a1, a2));
        dst[Ns] = cadd(m1, r1);
        dst[2 * Ns] = cadd(m2, r2);
        dst[3 * Ns] = csub(m2, r2);
        dst[4 * Ns] = csub(m1, r1);
      }
    }
  }
  __syncthreads();
}

// FFT of nl lines of length K = 2^a 3^b 5^c (nl*K <= TILE): radix-8 stages,
// one radix-4 or radix-2 stage for the rest of the power of two, radix-9
// stages and one radix-3 stage for the power of three, then radix 5.
template <typename C, int TILE, int TH = FFT_THREADS>
__device__ __forceinline__ void fft_lines(C* sh, int K, int pitch, int nl, int dir, const C* __restrict__ tw) {
  int Ns = 1, rest = K;
  while (rest % 8 == 0) {
    fft_stage<C, 8, (TILE + 8 * TH - 1) / (8 * TH)>(sh, K, pitch, nl, Ns, dir, tw);
    Ns *= 8;
    rest /= 8;
  }
  if (rest % 4 == 0) {
    fft_stage<C, 4, (TILE + 4 * TH - 1) / (4 * TH)>(sh, K, pitch, nl, Ns, dir, tw);
    Ns *= 4;
    rest /= 4;
  }
  if (rest % 2 == 0) {
    fft_stage<C, 2, (TILE + 2 * TH - 1) / (2 * TH)>(sh, K, pitch, nl, Ns, dir, tw);
    Ns *= 2;
    rest /= 2;
  }
  while (rest % 9 == 0) {
    fft_stage<C, 9, (TILE + 9 * TH - 1) / (9 * TH)>(sh, K, pitch, nl, Ns, dir, tw);
    Ns *= 9;
    rest /= 9;
  }
  if (rest % 3 == 0) {
    fft_stage<C, 3, (TILE + 3 * TH - 1) / (3 * TH)>(sh, K, pitch, nl, Ns, dir, tw);
    Ns *= 3;
    rest /= 3;
  }
  while (rest % 5 == 0) {
    fft_stage<C, 5, (TILE + 5 * TH - 1) / (5 * TH)>(sh, K, pitch, nl, Ns, dir, tw);
    Ns *= 5;
    rest /= 5;
  }
}

// ------------------------------------------------------------------- spreading
template <typename T, int P>
__device__ __forceinline__ void frac_weights(const RepDev& R, double x, double y, double z, int base[3],
                                             T W[3][P]) {
  double s2 = y * R.hinv[1];
  double s1 = (x - R.L2xp * s2) * R.hinv[0];
  double s3 = z * R.hinv[2];
  double s[3] = {s1, s2, s3};
  for (int d = 0; d < 3; ++d) {
    double u = s[d] * (double)R.K[d];
    double fl = floor(u);
    T fr = (T)(u - fl);
    int bi = ((int)fl) % R.K[d];
    base[d] = bi < 0 ? bi + R.K[d] : bi;
    bspline_weights<T, P>(fr, W[d]);
  }
}

// select W[j] for a runtime j without dynamic register indexing
template <typename T, int P> __device__ __forceinline__ T pick(const T* W, int j) {
  T w = (T)0;
#pragma unroll
  for (int k = 0; k < P; ++k) w = (k == j) ? W[k] : w;
  return w;
}

__device__ __forceinline__ int brick_of(const RepDev& R, const int* base) {
  return R.brick0 + ((base[2] / BRICK) * R.nbrick[1] + base[1] / BRICK) * R.nbrick[0] + base[0] / BRICK;
}

// Per bead: stencil base point and B-spline weights along each lattice vector,
// and a count in the brick holding the base point (for the gather).
template <typename T, int P>
__global__ void k_bead_prep(MobArgs a) {
  int s = a.slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.slot_begin + a.nslots) return;
  int r = a.slot_rep[s];
  if (r < 0) return;
  const RepDev& R = a.reps[r];
  if (!R.active || s - R.slot0 >= R.n) return;
  int base[3];
  T W[3][P];
  frac_weights<T, P>(R, a.rx[s], a.ry[s], a.rz[s], base, W);
  T* bw = (T*)a.bw + (size_t)BW_STRIDE * s;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a.bbase[3 * s + d] = base[d];
#pragma unroll
    for (int j = 0; j < P; ++j) bw[d * MAX_ORDER + j] = W[d][j];
  }
  if (a.p32) {
    // cluster-consistent FP32 position: the first bead of the cluster (a run of
    // <= 8 consecutive beads of one chain) plus the minimum image of this bead
    // relative to it, so a cluster is never split by the periodic wrap
    const int j = (s - R.slot0) % R.nb, f = s - (j & 7);
    double dx = a.rx[s] - a.rx[f], dy = a.ry[s] - a.ry[f], dz = a.rz[s] - a.rz[f];
    min_image_reduced(R, dx, dy, dz);
    a.p32[s] = make_float4((float)(a.rx[f] + dx), (float)(a.ry[f] + dy), (float)(a.rz[f] + dz), 0.f);
  }
  atomicAdd(&a.brick_cnt[brick_of(R, base)], 1);
}

// Per cluster: the largest distance of its beads from its first bead, stored in
// the first bead's FP32 position (.w), for the shared-image test of the cluster pass.
__global__ void k_cluster_prep(MobArgs a) {
  const int c = a.cl_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.cl_begin + a.ncl) return;
  const int first = a.cl_first[c], cnt = a.cl_cnt[c];
  const RepDev& R = a.reps[a.slot_rep[first]];
  if (!R.active) return;
  const float4 p0 = a.p32[first];
  float rad2 = 0.f;
  for (int k = 1; k < cnt; ++k) {
    const float4 q = a.p32[first + k];
    const float dx = q.x - p0.x, dy = q.y - p0.y, dz = q.z - p0.z;
    rad2 = fmaxf(rad2, dx * dx + dy * dy + dz * dz);
  }
  a.p32[first].w = sqrtf(rad2);
}

__global__ void k_brick_fill(MobArgs a) {
  int s = a.slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.slot_begin + a.nslots) return;
  int r = a.slot_rep[s];
  if (r < 0) return;
  const RepDev& R = a.reps[r];
  if (!R.active || s - R.slot0 >= R.n) return;
  int bi = brick_of(R, a.bbase + 3 * s);
  int k = atomicAdd(&a.brick_fill[bi], 1);
  a.brick_beads[a.brick_start[bi] + k] = s;
}

// Bead lists of the bricks in ascending slot order (k_brick_fill appends in
// arrival order), so the spreading sums do not depend on scheduling.
__global__ void k_brick_sort(MobArgs a) {
  int bi = a.brick_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (bi >= a.brick_begin + a.brick_len) return;
  int* L = a.brick_beads + a.brick_start[bi];
  const int n = a.brick_cnt[bi];
  for (int i = 1; i < n; ++i) {
    int v = L[i], k = i - 1;
    while (k >= 0 && L[k] > v) {
      L[k + 1] = L[k];
      --k;
    }
    L[k + 1] = v;
  }
}

// Lanczos normalisation vout = v / beta[r] (the gather and the real-space pass
// read vout) and, in mixed precision, the FP32 copy v32 of the scaled vector.
__global__ void k_scale_v(MobArgs a) {
  int s = a.slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.slot_begin + a.nslots) return;
  int r = a.slot_rep[s];
  if (r < 0) return;
  const RepDev& R = a.reps[r];
  if (!R.active || (a.done && a.done[r]) || s - R.slot0 >= R.n) return;
  double sc = a.beta ? (a.beta[r] > 0.0 ? 1.0 / a.beta[r] : 0.0) : 1.0;  // z = 0 gives y = 0
  double v0 = a.v[3 * s] * sc, v1 = a.v[3 * s + 1] * sc, v2 = a.v[3 * s + 2] * sc;
  if (a.vout) {
    a.vout[3 * s] = v0;
    a.vout[3 * s + 1] = v1;
    a.vout[3 * s + 2] = v2;
  }
  if (a.v32) a.v32[s] = make_float4((float)v0, (float)v1, (float)v2, 0.f);
}

// Spreading as a gather: one workgroup per 8x8x8 brick of one replica's grid,
// one thread per grid point.  A bead with stencil base b touches points b-P+1..b,
// so the beads reaching this brick live in it or in the next brick along each
// axis (P <= 8); they are staged through LDS in chunks, with stencil bases
// relative to the brick origin (in [0, K); grids are at least 16 long, so a
// base below the point never reaches it through the periodic wrap), and every
// thread adds the ones whose stencil covers its point.  Each wavefront is one z
// plane (8 x 8 points) and lists, per 4 x 4 quadrant of its lanes, the staged
// beads that reach the quadrant; a lane walks its quadrant's list, which holds
// about half the beads that reach the plane.  While listing, the wavefront
// stores each reaching bead's vector times its z weight at the plane and its
// packed x and y bases, so a lane's step reads that record and two weights.
// A point adds exactly the beads covering it, in staging order.  No atomics: the grid needs no zeroing and
// the result does not depend on scheduling.
#define GATHER_CHUNK 128
template <typename T, int P>
__global__ __launch_bounds__(512) void k_spread_gather(MobArgs a) {
  __shared__ int sb[GATHER_CHUNK * 3];
  __shared__ T sw[GATHER_CHUNK * 3 * P];
  __shared__ T sv[GATHER_CHUNK * 3];
  __shared__ unsigned char plist[BRICK][4][GATHER_CHUNK];  // per z plane and quadrant: staged beads reaching it
  // per z plane and staged bead: the vector times the bead's z weight at the
  // plane, and its x and y bases (packed), written by the plane's wavefront
  __shared__ T zv[BRICK][GATHER_CHUNK][3];
  __shared__ int zb[BRICK][GATHER_CHUNK];
  const BrickTask bt = a.bricks[blockIdx.x];
  const RepDev& R = a.reps[bt.rep];
  if (!R.active || (a.done && a.done[bt.rep])) return;
  const int tid = threadIdx.x, lane = tid & 63, wz = tid >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int K0 = R.K[0], K1 = R.K[1], K2 = R.K[2];
  const int g0 = bt.b0 * BRICK + (tid & 7), g1 = bt.b1 * BRICK + ((tid >> 3) & 7), g2 = bt.b2 * BRICK + (tid >> 6);
  const int lo0 = bt.b0 * BRICK, lo1 = bt.b1 * BRICK, lo2 = bt.b2 * BRICK;
  const int qd = ((lane >> 2) & 1) | (((lane >> 5) & 1) << 1);  // this lane's 4 x 4 quadrant
  // With scale_in_spread the input is normalised here (v / beta, as k_scale_v
  // would) and the bricks store the scaled input of their own beads (source
  // brick 0 is the brick itself, and every bead has one home brick).
  const bool fs = a.scale_in_spread != 0;
  const double* vin = (a.vout && !fs) ? a.vout : a.v;
  const double vsc = (fs && a.beta) ? (a.beta[bt.rep] > 0.0 ? 1.0 / a.beta[bt.rep] : 0.0) : 1.0;
  const T* bw = (const T*)a.bw;
  // the 8 source bricks (this one and its +1 neighbours): starts and counts in one step
  __shared__ int nbeg[8], noff[9];
  if (tid < 8) {
    int c0 = (bt.b0 + (tid & 1)) % R.nbrick[0];
    int c1 = (bt.b1 + ((tid >> 1) & 1)) % R.nbrick[1];
    int c2 = (bt.b2 + (tid >> 2)) % R.nbrick[2];
    int bi = R.brick0 + (c2 * R.nbrick[1] + c1) * R.nbrick[0] + c0;
    nbeg[tid] = a.brick_start[bi];
    noff[tid + 1] = a.brick_cnt[bi];
  }
  __syncthreads();
  if (tid == 0) {
    noff[0] = 0;
    for (int k = 0; k < 8; ++k) noff[k + 1] += noff[k];
  }
  __syncthreads();
  const int total = noff[8];
  T acc0 = 0, acc1 = 0, acc2 = 0;
  for (int k0 = 0; k0 < total; k0 += GATHER_CHUNK) {
    const int m = min(GATHER_CHUNK, total - k0);
    if (k0 > 0) __syncthreads();
    for (int k = tid; k < m * 3; k += blockDim.x) {
      int kk = k / 3, d = k - 3 * kk;
      int gi = k0 + kk, nbk = 0;
      while (gi >= noff[nbk + 1]) ++nbk;
      int s = a.brick_beads[nbeg[nbk] + gi - noff[nbk]];
      int rel = a.bbase[3 * s + d] - (d == 0 ? lo0 : d == 1 ? lo1 : lo2);
      rel += rel < 0 ? (d == 0 ? K0 : d == 1 ? K1 : K2) : 0;
      sb[k] = rel;
      if (fs) {
        const double x = vin[3 * s + d] * vsc;
        sv[k] = (T)x;
        if (nbk == 0) {
          if (a.vout) a.vout[3 * s + d] = x;
          if (a.v32) {
            float* v = (float*)(a.v32 + s);
            v[d] = (float)x;
            if (d == 0) v[3] = 0.f;
          }
        }
      } else if (a.v32) {
        const float4 v4 = a.v32[s];
        sv[k] = (T)(d == 0 ? v4.x : d == 1 ? v4.y : v4.z);
      } else {
        sv[k] = (T)vin[3 * s + d];
      }
#pragma unroll
      for (int j = 0; j < P; ++j) sw[k * P + j] = bw[(size_t)BW_STRIDE * s + d * MAX_ORDER + j];
    }
    __syncthreads();
    // list, in staging order, the beads that reach each quadrant of this plane
    // (ballot compaction): relative base b reaches points b-P+1..b, so a quadrant
    // starting at 4q is reached when 4q <= b <= 4q + 3 + P - 1
    int n0 = 0, n1 = 0, n2 = 0, n3 = 0;
    for (int k0c = 0; k0c < m; k0c += 64) {
      const int k = k0c + lane;
      bool c0 = false, c1 = false, c2 = false, c3 = false;
      if (k < m) {
        const int e0 = sb[3 * k], e1 = sb[3 * k + 1];
        const bool cz = (unsigned)(sb[3 * k + 2] - wz) < (unsigned)P;
        const bool x0 = e0 <= 2 + P, x1 = e0 >= 4 && e0 <= 6 + P;
        const bool y0 = e1 <= 2 + P, y1 = e1 >= 4 && e1 <= 6 + P;
        c0 = cz && x0 && y0;
        c1 = cz && x1 && y0;
        c2 = cz && x0 && y1;
        c3 = cz && x1 && y1;
        if (cz) {
          const T wzk = sw[(3 * k + 2) * P + (sb[3 * k + 2] - wz)];
          zv[wz][k][0] = wzk * sv[3 * k];
          zv[wz][k][1] = wzk * sv[3 * k + 1];
          zv[wz][k][2] = wzk * sv[3 * k + 2];
          zb[wz][k] = e0 | (e1 << 16);
        }
      }
      const unsigned long long b0 = __ballot(c0), b1 = __ballot(c1), b2 = __ballot(c2), b3 = __ballot(c3);
      if (c0) plist[wz][0][n0 + __popcll(b0 & below)] = (unsigned char)k;
      if (c1) plist[wz][1][n1 + __popcll(b1 & below)] = (unsigned char)k;
      if (c2) plist[wz][2][n2 + __popcll(b2 & below)] = (unsigned char)k;
      if (c3) plist[wz][3][n3 + __popcll(b3 & below)] = (unsigned char)k;
      n0 += __popcll(b0);
      n1 += __popcll(b1);
      n2 += __popcll(b2);
      n3 += __popcll(b3);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int mine = qd == 0 ? n0 : qd == 1 ? n1 : qd == 2 ? n2 : n3;
    const int gr0 = tid & 7, gr1 = (tid >> 3) & 7;
    const unsigned char* pl = plist[wz][qd];
    for (int q = 0; q < mine; ++q) {
      const int k = pl[q];
      const int e = zb[wz][k];
      const int d0 = (e & 0xffff) - gr0, d1 = (e >> 16) - gr1;
      if ((unsigned)d0 < (unsigned)P && (unsigned)d1 < (unsigned)P) {
        const T w = sw[(3 * k) * P + d0] * sw[(3 * k + 1) * P + d1];
        acc0 += w * zv[wz][k][0];
        acc1 += w * zv[wz][k][1];
        acc2 += w * zv[wz][k][2];
      }
    }
  }
  T* g = (T*)a.grid + R.grid0;
  const long long ncomp = (long long)K0 * K1 * K2;
  long long idx = ((long long)g2 * K1 + g1) * K0 + g0;
  g[idx] = acc0;
  g[idx + ncomp] = acc1;
  g[idx + 2 * ncomp] = acc2;
}

// Candidate beads of each brick for the matrix-core spread: the beads of the
// brick and of its +1 neighbours along each axis (the only ones whose stencil
// reaches it), concatenated in source-brick order, each list in slot order.
// Built once per position version; every bead is a candidate of exactly 8
// bricks, so the lists hold 8 entries per bead.
__global__ void k_brick_cand_count(MobArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.brick_len) return;
  const BrickTask bt = a.bricks[j];
  const RepDev& R = a.reps[bt.rep];
  int n = 0;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int c0 = (bt.b0 + (t & 1)) % R.nbrick[0], c1 = (bt.b1 + ((t >> 1) & 1)) % R.nbrick[1];
    const int c2 = (bt.b2 + (t >> 2)) % R.nbrick[2];
    n += a.brick_cnt[R.brick0 + (c2 * R.nbrick[1] + c1) * R.nbrick[0] + c0];
  }
  a.bcand_cnt[a.brick_begin + j] = n;
}

__global__ __launch_bounds__(64) void k_brick_cand_fill(MobArgs a) {
  const int j = blockIdx.x, bi = a.brick_begin + j;
  const BrickTask bt = a.bricks[j];
  const RepDev& R = a.reps[bt.rep];
  int* dst = a.bcand + a.bcand_start[bi];
  for (int t = 0; t < 8; ++t) {
    const int c0 = (bt.b0 + (t & 1)) % R.nbrick[0], c1 = (bt.b1 + ((t >> 1) & 1)) % R.nbrick[1];
    const int c2 = (bt.b2 + (t >> 2)) % R.nbrick[2];
    const int sb = R.brick0 + (c2 * R.nbrick[1] + c1) * R.nbrick[0] + c0;
    const int n = a.brick_cnt[sb], src = a.brick_start[sb];
    for (int i = threadIdx.x; i < n; i += 64) dst[i] = a.brick_beads[src + i];
    dst += n;
  }
}

// One bead's B-spline weights along one axis at the 8 points of a brick:
// out[x] = W[e - x] where 0 <= e - x < P, else 0, for a relative base
// e = 8h + f in [0, 16).  W[(f - x) mod 8] is a rotation of the reversed weights
// by f: three conditional rotations as bitwise selects under a mask the compiler
// treats as opaque (left to itself it turns the rotation into a compare chain
// per output), kept where x <= f for h = 0 and where x > f for h = 1.
template <int P>
__device__ __forceinline__ void spm_expand(const float4 w0, const float4 w1, int e, float out[8]) {
  const float W[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
  unsigned u[8];
#pragma unroll
  for (int y = 0; y < 8; ++y) {
    const int j = (8 - y) & 7;
    u[y] = j < P ? __float_as_uint(W[j]) : 0u;  // u[y] = W[-y mod 8]; weights past the order are zero
  }
  const int f = e & 7, h = e >> 3;
#pragma unroll
  for (int b = 1; b < 8; b <<= 1) {  // rotate right by f: u[y] <- u[(y - f) mod 8] = W[(f - y) mod 8]
    unsigned m = 0u - (unsigned)((f & b) != 0);
    asm volatile("" : "+v"(m));
    unsigned t[8];
#pragma unroll
    for (int y = 0; y < 8; ++y) t[y] = (u[(y + 8 - b) & 7] & m) | (u[y] & ~m);
#pragma unroll
    for (int y = 0; y < 8; ++y) u[y] = t[y];
  }
  // kept points: x <= f for h = 0, x > f for h = 1
  const unsigned low = (2u << f) - 1u, keep = h == 0 ? low : ~low;
#pragma unroll
  for (int x = 0; x < 8; ++x) out[x] = __uint_as_float(u[x] & (0u - ((keep >> x) & 1u)));
}

// Spreading on the matrix cores (FP32 grids).  A brick's grid values are a sum of
// separable terms, G[x,y,z,c] = sum_b wx_b(x) wy_b(y) wz_b(z) v_b(c), i.e. the
// product of a 64 x K matrix A[(x,y), b] = wx_b(x) wy_b(y) and a K x 24 matrix
// B[b, (z,c)] = wz_b(z) v_b(c) over the K candidate beads of the brick.  One
// workgroup per brick, four wavefronts; wavefront w owns the 16 rows x = 0..7,
// y = 2w, 2w+1 and two 16-column tiles of the product (n = 3z + c < 24; the
// second tile's columns 24-31 read whatever follows a row of B and are never
// stored), and accumulates them with v_mfma_f32_16x16x4f32 over the beads four
// at a time.  Beads are staged in chunks by two threads each (x and y weights;
// the row of B), which load the bead's base, weights and vector and write its
// expanded rows to LDS (20 KB per workgroup, so seven workgroups fit a CU and
// hide each other's staging latency).  The sum
// runs over the beads in staging order with the matrix cores' fixed order, so the
// result does not depend on scheduling; the output tile goes through LDS so the
// stores are runs of 8 consecutive points as in the gather.
#define SPM_CHUNK 128
#define SPM_OUT_PITCH 65
typedef float spm_f32x4 __attribute__((ext_vector_type(4)));
template <int P>
__global__ __launch_bounds__(256) void k_spread_mfma(MobArgs a) {
  __shared__ float4 sE[SPM_CHUNK][4];  // x weights at x = 0..7, then y weights at y = 0..7
  // rows of B (n = 0..23), then 8 floats that the last row's reads of the second
  // column tile run into (columns n >= 24 of the product are never stored); reused
  // for the output tile
  __shared__ float4 sB[SPM_CHUNK * 6 + 2];
  static_assert(sizeof(sB) >= 3 * BRICK * SPM_OUT_PITCH * sizeof(float), "output tile fits the B stage");
  static_assert(SPM_CHUNK == 128, "staging maps the chunk's 128 beads to each half of the workgroup");
  // XCD-aware order: each XCD takes a contiguous range of bricks, so the neighbours
  // that share a bead's data, and that complete each other's cache lines of the
  // grid, run under one L2
  const int jb = xcd_block(blockIdx.x, gridDim.x), bi = a.brick_begin + jb;
  const BrickTask bt = a.bricks[jb];
  const int cb = a.bcand_start[bi], total = a.bcand_cnt[bi], nhome = a.brick_cnt[bi];
  const RepDev& R = a.reps[bt.rep];
  if (!R.active || (a.done && a.done[bt.rep])) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int K0 = R.K[0], K1 = R.K[1], K2 = R.K[2];
  const int lo[3] = {bt.b0 * BRICK, bt.b1 * BRICK, bt.b2 * BRICK};
  const bool fs = a.scale_in_spread != 0;
  const double* vin = (a.vout && !fs) ? a.vout : a.v;
  const double vsc = (fs && a.beta) ? (a.beta[bt.rep] > 0.0 ? 1.0 / a.beta[bt.rep] : 0.0) : 1.0;
  const float* bw = (const float*)a.bw;
  // this lane's operands: row (ax, ay) of A, column bj of both tiles of B, bead kq of each group of four
  const int ax = lane & 7, ay = 2 * w + ((lane >> 3) & 1), kq = lane >> 4, bj = lane & 15;
  const float* sEf = (const float*)sE;
  const float* sBf = (const float*)sB;
  spm_f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < total; k0 += SPM_CHUNK) {
    const int m = min(SPM_CHUNK, total - k0), mp = (m + 15) & ~15;
    if (k0 > 0) __syncthreads();  // the previous chunk's products have read the stage
    // bead k = tid % 128 of the chunk: wavefronts 0-1 stage its x and y weights,
    // wavefronts 2-3 its row of B (z weights times the vector); each thread issues
    // all its loads at once; rows m <= k < mp (padding to whole groups of 16) are zero
    {
      const int k = tid & 127, half = tid >> 7;
      if (k < mp && k >= m) {
        const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
        float4* dst = half == 0 ? sE[k] : &sB[6 * k];
        for (int q = 0; q < (half == 0 ? 4 : 6); ++q) dst[q] = zero;
      } else if (k < m) {
        const int kk = k0 + k, s = a.bcand[cb + kk];
        const float4* wp = (const float4*)(bw + (size_t)BW_STRIDE * s);
        float o[8];
        if (half == 0) {
          int ex = a.bbase[3 * s] - lo[0], ey = a.bbase[3 * s + 1] - lo[1];
          const float4 x0 = wp[0], x1 = wp[1], y0 = wp[2], y1 = wp[3];
          ex += ex < 0 ? K0 : 0;
          ey += ey < 0 ? K1 : 0;
          spm_expand<P>(x0, x1, ex, o);
          sE[k][0] = make_float4(o[0], o[1], o[2], o[3]);
          sE[k][1] = make_float4(o[4], o[5], o[6], o[7]);
          spm_expand<P>(y0, y1, ey, o);
          sE[k][2] = make_float4(o[0], o[1], o[2], o[3]);
          sE[k][3] = make_float4(o[4], o[5], o[6], o[7]);
        } else {
          int ez = a.bbase[3 * s + 2] - lo[2];
          const float4 z0 = wp[2 * MAX_ORDER / 4], z1 = wp[2 * MAX_ORDER / 4 + 1];
          ez += ez < 0 ? K2 : 0;
          float v0, v1, v2;
          if (fs) {
            const double x0 = vin[3 * s] * vsc, x1 = vin[3 * s + 1] * vsc, x2 = vin[3 * s + 2] * vsc;
            v0 = (float)x0;
            v1 = (float)x1;
            v2 = (float)x2;
            if (kk < nhome) {  // every bead has one home brick: it stores the scaled input once
              if (a.vout) {
                a.vout[3 * s] = x0;
                a.vout[3 * s + 1] = x1;
                a.vout[3 * s + 2] = x2;
              }
              if (a.v32) a.v32[s] = make_float4(v0, v1, v2, 0.f);
            }
          } else if (a.v32) {
            const float4 v4 = a.v32[s];
            v0 = v4.x;
            v1 = v4.y;
            v2 = v4.z;
          } else {
            v0 = (float)vin[3 * s];
            v1 = (float)vin[3 * s + 1];
            v2 = (float)vin[3 * s + 2];
          }
          spm_expand<P>(z0, z1, ez, o);
          // row of B: n = 3z + c -> o[z] * v[c], n = 0..23
          float4* dst = &sB[6 * k];
          dst[0] = make_float4(o[0] * v0, o[0] * v1, o[0] * v2, o[1] * v0);
          dst[1] = make_float4(o[1] * v1, o[1] * v2, o[2] * v0, o[2] * v1);
          dst[2] = make_float4(o[2] * v2, o[3] * v0, o[3] * v1, o[3] * v2);
          dst[3] = make_float4(o[4] * v0, o[4] * v1, o[4] * v2, o[5] * v0);
          dst[4] = make_float4(o[5] * v1, o[5] * v2, o[6] * v0, o[6] * v1);
          dst[5] = make_float4(o[6] * v2, o[7] * v0, o[7] * v1, o[7] * v2);
        }
      }
    }
    __syncthreads();
    for (int kk = 0; kk < mp; kk += 16) {  // four groups of four beads: operands first, then the products
      float av[4], b0[4], b1[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kb = kk + 4 * u + kq;
        av[u] = sEf[16 * kb + ax] * sEf[16 * kb + 8 + ay];
        b0[u] = sBf[24 * kb + bj];
        b1[u] = sBf[24 * kb + 16 + bj];  // bj >= 8: the next row's columns, in unused output columns
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], b0[u], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], b1[u], acc1, 0, 0, 0);
      }
    }
  }
  __syncthreads();
  // D[i][j] sits in register i % 4 of lane j + 16 (i / 4): row i = (x, y), column (z, c)
  // out[(c * 8 + z) * SPM_OUT_PITCH + y * 8 + x]: the pitch of 65 puts the 16 (z, c) of one
  // register on distinct banks
  float* out = (float*)sB;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * kq + r, x = i & 7, y = 2 * w + (i >> 3);
    {
      const int n = bj, z = n / 3, c = n - 3 * z;
      out[(c * BRICK + z) * SPM_OUT_PITCH + y * BRICK + x] = acc0[r];
    }
    if (bj < 3 * BRICK - 16) {
      const int n = 16 + bj, z = n / 3, c = n - 3 * z;
      out[(c * BRICK + z) * SPM_OUT_PITCH + y * BRICK + x] = acc1[r];
    }
  }
  __syncthreads();
  float* g = (float*)a.grid + R.grid0;
  const long long ncomp = (long long)K0 * K1 * K2;
  for (int idx = tid; idx < 3 * BRICK * BRICK * BRICK; idx += 256) {
    const int cz = idx >> 6, y = (idx >> 3) & 7, x = idx & 7, c = cz >> 3, z = cz & 7;
    g[c * ncomp + ((long long)(lo[2] + z) * K1 + lo[1] + y) * K0 + lo[0] + x] = out[cz * SPM_OUT_PITCH + (idx & 63)];
  }
}

// ------------------------------------------------------------------- FFT passes
// Copies a line length's twiddle table into LDS, so the Stockham stages read
// their twiddles there; the barrier that follows the grid load orders it.
template <typename C> __device__ __forceinline__ void stage_twiddles(C* dst, const C* __restrict__ src, int K) {
  for (int i = threadIdx.x; i < K; i += blockDim.x) dst[i] = src[i];
}

// x pass: task = (rep, comp, first row pair, pairs); forward packs rows 2p, 2p+1
// as one complex line and splits the spectrum into two half spectra (Nyquist
// dropped); inverse rebuilds the full line and writes real/imag to the rows.
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_x(MobArgs a, int dir) {
  typedef typename Prec<T>::C C;
  constexpr int TILE = sizeof(T) == 4 ? FFT_TILE_X : FFT_TILE_X64;
  __shared__ C sh[TILE + FFT_MAXL];
  __shared__ C tw[KMAX];
  FftTask tk = a.tx[blockIdx.x];
  const RepDev& R = a.reps[tk.rep];
  if (!R.active || (a.done && a.done[tk.rep])) return;
  const int K = R.K[0], Kh = K / 2, pitch = K + 1, nl = tk.nl;
  const FDiv dK(K), dKh(Kh);
  const long long ncomp = (long long)R.K[0] * R.K[1] * R.K[2];
  T* g = (T*)a.grid + R.grid0 + tk.comp * ncomp;
  C* gc = (C*)g;
  stage_twiddles(tw, (const C*)a.tw + R.twoff[0], K);
  if (dir < 0) {
    for (int e = threadIdx.x; e < nl * K; e += blockDim.x) {
      int l = dK.div(e), x = e - l * K;
      long long row = 2LL * (tk.l0 + l);
      C z;
      z.x = g[row * K + x];
      z.y = g[(row + 1) * K + x];
      sh[l * pitch + x] = z;
    }
    __syncthreads();
    fft_lines<C, TILE>(sh, K, pitch, nl, -1, tw);
    for (int e = threadIdx.x; e < nl * Kh; e += blockDim.x) {
      int l = dKh.div(e), k = e - l * Kh;
      C zk = sh[l * pitch + k], zm = sh[l * pitch + (k == 0 ? 0 : K - k)];
      C A, B;
      A.x = (T)0.5 * (zk.x + zm.x);
      A.y = (T)0.5 * (zk.y - zm.y);
      B.x = (T)0.5 * (zk.y + zm.y);
      B.y = (T)0.5 * (zm.x - zk.x);
      long long row = 2LL * (tk.l0 + l);
      gc[row * Kh + k] = A;
      gc[(row + 1) * Kh + k] = B;
    }
  } else {
    for (int e = threadIdx.x; e < nl * Kh; e += blockDim.x) {
      int l = dKh.div(e), k = e - l * Kh;
      long long row = 2LL * (tk.l0 + l);
      C A = gc[row * Kh + k], B = gc[(row + 1) * Kh + k];
      C z;
      z.x = A.x - B.y;
      z.y = A.y + B.x;
      sh[l * pitch + k] = z;
      if (k > 0) {
        C w;
        w.x = A.x + B.y;
        w.y = B.x - A.y;
        sh[l * pitch + K - k] = w;
      } else {
        C zero;
        zero.x = (T)0;
        zero.y = (T)0;
        sh[l * pitch + Kh] = zero;
      }
    }
    __syncthreads();
    fft_lines<C, TILE>(sh, K, pitch, nl, 1, tw);
    for (int e = threadIdx.x; e < nl * K; e += blockDim.x) {
      int l = dK.div(e), x = e - l * K;
      long long row = 2LL * (tk.l0 + l);
      C z = sh[l * pitch + x];
      g[row * K + x] = z.x;
      g[(row + 1) * K + x] = z.y;
    }
  }
}

// Fused x-y pass over one z plane of one component (planes with K0*K1/2 <=
// XY_ELEMS).  Forward: the K1 real rows are packed in pairs as K1/2 complex
// lines and transformed along x; each pair is split into the two rows' half
// spectra (Nyquist dropped), which are written transposed, one line per x
// wavenumber, and transformed along y; the result goes out in the layout the
// y pass writes, (z*K1 + y)*Kh + x.  Inverse: the same steps backwards.  The
// two LDS layouts share one buffer: the split and the rebuild go through
// registers, all reads before a barrier, all writes after it.
template <typename T>
__global__ __launch_bounds__(XY_THREADS) void k_fft_xy(MobArgs a, int dir) {
  typedef typename Prec<T>::C C;
  constexpr int MAXP = (XY_ELEMS / 2 + XY_THREADS - 1) / XY_THREADS;  // row pairs x wavenumbers per thread
  __shared__ C sh[XY_ELEMS + KMAX / 2];
  __shared__ C twx[KMAX], twy[KMAX];
  FftTask tk = a.txy[blockIdx.x];
  const RepDev& R = a.reps[tk.rep];
  if (!R.active || (a.done && a.done[tk.rep])) return;
  const int K0 = R.K[0], K1 = R.K[1], Kh = K0 / 2, P2 = K1 / 2;
  const FDiv dK0(K0), dKh(Kh);
  const int pa = K0 + 1, pb = K1 + 1;  // x-line and y-line pitches
  const long long ncomp = (long long)K0 * K1 * R.K[2];
  const int z = tk.l0;
  T* g = (T*)a.grid + R.grid0 + tk.comp * ncomp + (long long)z * K1 * K0;  // this plane's real rows
  C* gc = (C*)((T*)a.grid + R.grid0 + tk.comp * ncomp) + (long long)z * K1 * Kh;  // its half spectrum
  stage_twiddles(twx, (const C*)a.tw + R.twoff[0], K0);
  stage_twiddles(twy, (const C*)a.tw + R.twoff[1], K1);
  const int t = threadIdx.x, TH = blockDim.x, np = P2 * Kh;
  C ra[MAXP], rb[MAXP];
  if (dir < 0) {
    for (int e = t; e < P2 * K0; e += TH) {
      int p = dK0.div(e), x = e - p * K0;
      C v;
      v.x = g[(2 * p) * K0 + x];
      v.y = g[(2 * p + 1) * K0 + x];
      sh[p * pa + x] = v;
    }
    __syncthreads();
    fft_lines<C, XY_ELEMS, XY_THREADS>(sh, K0, pa, P2, -1, twx);
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      int e = t + q * TH;
      if (e < np) {
        int p = dKh.div(e), k = e - p * Kh;
        C zk = sh[p * pa + k], zm = sh[p * pa + (k == 0 ? 0 : K0 - k)];
        ra[q].x = (T)0.5 * (zk.x + zm.x);
        ra[q].y = (T)0.5 * (zk.y - zm.y);
        rb[q].x = (T)0.5 * (zk.y + zm.y);
        rb[q].y = (T)0.5 * (zm.x - zk.x);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      int e = t + q * TH;
      if (e < np) {
        int p = dKh.div(e), k = e - p * Kh;
        sh[k * pb + 2 * p] = ra[q];
        sh[k * pb + 2 * p + 1] = rb[q];
      }
    }
    __syncthreads();
    fft_lines<C, XY_ELEMS, XY_THREADS>(sh, K1, pb, Kh, -1, twy);
    for (int e = t; e < K1 * Kh; e += TH) {
      int y = dKh.div(e), k = e - y * Kh;
      gc[(long long)y * Kh + k] = sh[k * pb + y];
    }
  } else {
    for (int e = t; e < K1 * Kh; e += TH) {
      int y = dKh.div(e), k = e - y * Kh;
      sh[k * pb + y] = gc[(long long)y * Kh + k];
    }
    __syncthreads();
    fft_lines<C, XY_ELEMS, XY_THREADS>(sh, K1, pb, Kh, 1, twy);
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      int e = t + q * TH;
      if (e < np) {
        int p = dKh.div(e), k = e - p * Kh;
        ra[q] = sh[k * pb + 2 * p];
        rb[q] = sh[k * pb + 2 * p + 1];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      int e = t + q * TH;
      if (e < np) {
        int p = dKh.div(e), k = e - p * Kh;
        C A = ra[q], B = rb[q], v;
        v.x = A.x - B.y;
        v.y = A.y + B.x;
        sh[p * pa + k] = v;
        if (k > 0) {
          C w;
          w.x = A.x + B.y;
          w.y = B.x - A.y;
          sh[p * pa + K0 - k] = w;
        } else {
          C zero;
          zero.x = (T)0;
          zero.y = (T)0;
          sh[p * pa + Kh] = zero;
        }
      }
    }
    __syncthreads();
    fft_lines<C, XY_ELEMS, XY_THREADS>(sh, K0, pa, P2, 1, twx);
    for (int e = t; e < P2 * K0; e += TH) {
      int p = dK0.div(e), x = e - p * K0;
      C v = sh[p * pa + x];
      g[(2 * p) * K0 + x] = v.x;
      g[(2 * p + 1) * K0 + x] = v.y;
    }
  }
}

// y pass (C2C along a2): line L = z*Kh + x, element k at (z*K1 + k)*Kh + x.
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_y(MobArgs a, int dir) {
  typedef typename Prec<T>::C C;
  constexpr int TILE = sizeof(T) == 4 ? FFT_TILE_Y : FFT_TILE_Y64;
  __shared__ C sh[TILE + FFT_MAXL];
  __shared__ C tw[KMAX];
  FftTask tk = a.ty[blockIdx.x];
  const RepDev& R = a.reps[tk.rep];
  if (!R.active || (a.done && a.done[tk.rep])) return;
  const int Kh = R.K[0] / 2, K = R.K[1], pitch = K + 1, nl = tk.nl;
  const FDiv dnl(nl), dKh(Kh);
  const long long ncomp = (long long)R.K[0] * R.K[1] * R.K[2];
  C* gc = (C*)((T*)a.grid + R.grid0 + tk.comp * ncomp);
  stage_twiddles(tw, (const C*)a.tw + R.twoff[1], K);
  for (int e = threadIdx.x; e < nl * K; e += blockDim.x) {
    int k = dnl.div(e), l = e - k * nl;
    int L = tk.l0 + l, z = dKh.div(L), x = L - z * Kh;
    sh[l * pitch + k] = gc[((long long)z * K + k) * Kh + x];
  }
  __syncthreads();
  fft_lines<C, TILE>(sh, K, pitch, nl, dir, tw);
  for (int e = threadIdx.x; e < nl * K; e += blockDim.x) {
    int k = dnl.div(e), l = e - k * nl;
    int L = tk.l0 + l, z = dKh.div(L), x = L - z * Kh;
    gc[((long long)z * K + k) * Kh + x] = sh[l * pitch + k];
  }
}

// z pass: forward FFT of the three components, Green multiply, inverse FFT.
// Line L = y*Kh + x, element k at k*(K1*Kh) + L.
template <typename T, int TILE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_z_green(MobArgs a) {
  typedef typename Prec<T>::C C;
  __shared__ C sh[3 * (TILE + FFT_MAXL)];
  __shared__ C tw[KMAX];
  FftTask tk = a.tz[blockIdx.x];
  const RepDev& R = a.reps[tk.rep];
  if (!R.active || (a.done && a.done[tk.rep])) return;
  stage_twiddles(tw, (const C*)a.tw + R.twoff[2], R.K[2]);
  const int K0 = R.K[0], Kh = K0 / 2, K1 = R.K[1], K = R.K[2], pitch = K + 1, nl = tk.nl;
  const FDiv dnl(nl), dKh(Kh);
  const long long ncomp = (long long)R.K[0] * R.K[1] * R.K[2];
  const long long stride = (long long)K1 * Kh;
  C* gc[3];
  for (int c = 0; c < 3; ++c) gc[c] = (C*)((T*)a.grid + R.grid0 + c * ncomp);
  const int cpitch = nl * pitch;
  for (int e = threadIdx.x; e < nl * K; e += blockDim.x) {
    int k = dnl.div(e), l = e - k * nl;
    long long idx = k * stride + tk.l0 + l;
    for (int c = 0; c < 3; ++c) sh[c * cpitch + l * pitch + k] = gc[c][idx];
  }
  __syncthreads();
  fft_lines<C, 3 * TILE>(sh, K, pitch, 3 * nl, -1, tw);
  const T* bsp = (const T*)a.bsp + R.bsp0;
  const T ia2 = (T)(1.0 / (4.0 * R.alpha * R.alpha));
  const T invV = (T)(1.0 / R.vol);
  for (int e = threadIdx.x; e < nl * K; e += blockDim.x) {
    int k = dnl.div(e), l = e - k * nl;
    int L = tk.l0 + l, y = dKh.div(L), x = L - y * Kh;
    C* p0 = sh + l * pitch + k;
    C* p1 = p0 + cpitch;
    C* p2 = p1 + cpitch;
    bool zero = (y == K1 / 2) || (k == K / 2) || (x == 0 && y == 0 && k == 0);
    if (zero) {
      C z0;
      z0.x = (T)0;
      z0.y = (T)0;
      *p0 = z0;
      *p1 = z0;
      *p2 = z0;
      continue;
    }
    int m1 = x, m2 = (y < K1 / 2) ? y : y - K1, m3 = (k < K / 2) ? k : k - K;
    T kx = (T)m1 * (T)R.kb[0];
    T ky = (T)m1 * (T)R.kb[1] + (T)m2 * (T)R.kb[2];
    T kz = (T)m3 * (T)R.kb[3];
    T kk = kx * kx + ky * ky + kz * kz;
    T q = kk * ia2;
    T ikk = (T)1 / kk;  // the one division per wave vector
    T gsc = ((T)1 - kk * (T)(1.0 / 3.0)) * ((T)1 + q + (T)2 * q * q) * ((T)18.849555921538759 * ikk) * exp(-q) * invV;
    gsc *= bsp[x] * bsp[K0 + y] * bsp[K0 + K1 + k];
    C v0 = *p0, v1 = *p1, v2 = *p2;
    T dr = (kx * v0.x + ky * v1.x + kz * v2.x) * ikk;
    T di = (kx * v0.y + ky * v1.y + kz * v2.y) * ikk;
    C o0, o1, o2;
    o0.x = gsc * (v0.x - kx * dr);
    o0.y = gsc * (v0.y - kx * di);
    o1.x = gsc * (v1.x - ky * dr);
    o1.y = gsc * (v1.y - ky * di);
    o2.x = gsc * (v2.x - kz * dr);
    o2.y = gsc * (v2.y - kz * di);
    *p0 = o0;
    *p1 = o1;
    *p2 = o2;
  }
  __syncthreads();
  fft_lines<C, 3 * TILE>(sh, K, pitch, 3 * nl, 1, tw);
  for (int e = threadIdx.x; e < nl * K; e += blockDim.x) {
    int k = dnl.div(e), l = e - k * nl;
    long long idx = k * stride + tk.l0 + l;
    for (int c = 0; c < 3; ++c) gc[c][idx] = sh[c * cpitch + l * pitch + k];
  }
}

// -------------------------------------------------- real space + interpolation
// Interpolation of one bead's velocity from the three component grids by one
// wavefront.  Lane j1 + P*j2 (< P^2) owns stencil column (j1, j2) and walks the
// P planes j3, so each load instruction reads P runs of P consecutive x points
// (P rows of one plane) rather than P^2 scattered rows.
template <typename T, int P>
__device__ __forceinline__ void interp_bead(const MobArgs& a, const RepDev& R, int s, int lane, CompSum<T> acc[3]) {
  s = __builtin_amdgcn_readfirstlane(s);  // wave-uniform: the stencil base and weights are scalar loads
  if (lane >= P * P) return;
  int base[3];
  T W[3][P];
  const T* bw = (const T*)a.bw + (size_t)BW_STRIDE * s;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    base[d] = a.bbase[3 * s + d];
#pragma unroll
    for (int j = 0; j < P; ++j) W[d][j] = bw[d * MAX_ORDER + j];
  }
  const int j1 = lane % P, j2 = lane / P;
  const int K0 = R.K[0], K1 = R.K[1], K2 = R.K[2];
  const long long plane = (long long)K1 * K0, ncomp = plane * K2;
  int g1 = base[0] - j1, g2 = base[1] - j2;
  g1 += g1 < 0 ? K0 : 0;
  g2 += g2 < 0 ? K1 : 0;
  const T w12 = pick<T, P>(W[0], j1) * pick<T, P>(W[1], j2);
  T s0 = 0, s1 = 0, s2 = 0;
  if constexpr (sizeof(T) == 4) {
    // buffer loads: the lane's column as a 32-bit byte offset, the plane and
    // component as a scalar one (a replica's grid is below 2 GB)
    const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc(
        (void*)((const T*)a.grid + R.grid0), (short)0, (int)(3 * ncomp * (long long)sizeof(T)), 0x00020000);
    const int vo = (g2 * K0 + g1) * (int)sizeof(T);
#pragma unroll
    for (int j3 = 0; j3 < P; ++j3) {
      int g3 = base[2] - j3;
      g3 += g3 < 0 ? K2 : 0;
      const int so = (int)(g3 * plane * (long long)sizeof(T));
      const int sc = (int)(ncomp * (long long)sizeof(T));
      const T wt = W[2][j3];
      s0 += wt * __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(rg, vo, so, 0));
      s1 += wt * __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(rg, vo, so + sc, 0));
      s2 += wt * __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(rg, vo, so + 2 * sc, 0));
    }
  } else {
    const T* col = (const T*)a.grid + R.grid0 + (long long)g2 * K0 + g1;
#pragma unroll
    for (int j3 = 0; j3 < P; ++j3) {
      int g3 = base[2] - j3;
      g3 += g3 < 0 ? K2 : 0;
      const T* q = col + g3 * plane;
      const T wt = W[2][j3];
      s0 += wt * q[0];
      s1 += wt * q[ncomp];
      s2 += wt * q[2 * ncomp];
    }
  }
  acc[0].add(w12 * s0);
  acc[1].add(w12 * s1);
  acc[2].add(w12 * s2);
}

// One wavefront per bead.  Lanes split the interpolation stencil rows and the
// real-space list; each lane keeps compensated FP32 (or FP64) sums, which are
// combined across the wavefront in FP64.
#ifndef PAIR_UNROLL
#define PAIR_UNROLL 4
#endif
template <typename T, int P>
__global__ __launch_bounds__(256) void k_mob_final(MobArgs a) {
  const int lane = threadIdx.x & 63;
  int s = a.slot_begin + xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (s >= a.slot_begin + a.nslots) return;
  int r = a.slot_rep[s];
  if (r < 0) return;
  const RepDev& R = a.reps[r];
  if (!R.active || (a.done && a.done[r])) return;
  if (s - R.slot0 >= R.n) {
    if (lane == 0) a.u[3 * s] = a.u[3 * s + 1] = a.u[3 * s + 2] = 0.0;
    return;
  }
  const double* vv = a.vout ? a.vout : a.v;
  const double sc = (a.beta && !a.vout) ? (a.beta[r] > 0.0 ? 1.0 / a.beta[r] : 0.0) : 1.0;
  const double xi = a.rx[s], yi = a.ry[s], zi = a.rz[s];
  CompSum<T> acc[3];
  interp_bead<T, P>(a, R, s, lane, acc);  // reciprocal part
  // real-space part over the list
  const T alpha = (T)R.alpha, rc2 = (T)R.rc2;
  const int m = a.nrpy[s];
  if constexpr (sizeof(T) == 4) {  // mixed precision: FP32 positions and vector, one 16-byte load each per partner
    const float4 pi = a.p32[s];
    const float rd0 = (float)R.rd[0], rd1 = (float)R.rd[1], rd2 = (float)R.rd[2], rd3 = (float)R.rd[3];
    const float rb0 = (float)R.rb[0], rb1 = (float)R.rb[1], rb2 = (float)R.rb[2], rb3 = (float)R.rb[3];
    const float Lz = (float)R.L, iLz = (float)R.invL, rcf = (float)R.rc2;
    const int* row = a.lrpy + (size_t)s * a.cap_rpy;
    bool overlap = false;
    // PAIR_UNROLL partners per lane in flight: all loads issued before any arithmetic
    for (int k0 = lane; k0 < m; k0 += 64 * PAIR_UNROLL) {
      int jj[PAIR_UNROLL];
      float4 pj[PAIR_UNROLL], vj[PAIR_UNROLL];
#pragma unroll
      for (int u = 0; u < PAIR_UNROLL; ++u) jj[u] = (k0 + 64 * u < m) ? row[k0 + 64 * u] : -1;
#pragma unroll
      for (int u = 0; u < PAIR_UNROLL; ++u) {
        const int j = jj[u] < 0 ? s : jj[u];
        pj[u] = a.p32[j];
        vj[u] = a.v32[j];
      }
#pragma unroll
      for (int u = 0; u < PAIR_UNROLL; ++u) {
        if (jj[u] < 0) continue;
        const int j = jj[u];
        float dx = pj[u].x - pi.x, dy = pj[u].y - pi.y, dz = pj[u].z - pi.z;
        const float n1 = rintf(rd0 * dx + rd1 * dy), n2 = rintf(rd2 * dx + rd3 * dy);
        dx -= n1 * rb0 + n2 * rb2;
        dy -= n1 * rb1 + n2 * rb3;
        dz -= Lz * rintf(dz * iLz);
        const float rr = dx * dx + dy * dy + dz * dz;
        if (rr < RTAB_R0 * RTAB_R0) {
          overlap = true;  // evaluated below in FP64
        } else if (rr < rcf) {
          // t1 and t2/r^2 from the replica's cubic table (FP64-fitted, ~1e-7 relative)
          int it;
          float t;
          rtab_locate(rr, it, t);
          const float4 ca = a.rtab[R.tab0 + 2 * it], cb = a.rtab[R.tab0 + 2 * it + 1];
          const float t1 = ca.x + t * (ca.y + t * (ca.z + t * ca.w));
          const float b2 = cb.x + t * (cb.y + t * (cb.z + t * cb.w));
          const float dv = (dx * vj[u].x + dy * vj[u].y + dz * vj[u].z) * b2;
          acc[0].add((T)(t1 * vj[u].x + dv * dx));
          acc[1].add((T)(t1 * vj[u].y + dv * dy));
          acc[2].add((T)(t1 * vj[u].z + dv * dz));
        }
      }
    }
    // pairs closer than RTAB_R0 (rare in a physical configuration): the overlap
    // correction cancels 1/r^3 terms, so they are evaluated in FP64 from the FP64
    // positions, in a second pass only by the wavefronts that met one
    if (__ballot(overlap)) {
      for (int k = lane; k < m; k += 64) {
        const int j = row[k];
        double ex = a.rx[j] - xi, ey = a.ry[j] - yi, ez = a.rz[j] - zi;
        float fx = a.p32[j].x - pi.x, fy = a.p32[j].y - pi.y, fz = a.p32[j].z - pi.z;
        const float n1 = rintf(rd0 * fx + rd1 * fy), n2 = rintf(rd2 * fx + rd3 * fy);
        fx -= n1 * rb0 + n2 * rb2;
        fy -= n1 * rb1 + n2 * rb3;
        fz -= Lz * rintf(fz * iLz);
        if (!(fx * fx + fy * fy + fz * fz < RTAB_R0 * RTAB_R0)) continue;  // the same test as the first pass
        min_image_reduced(R, ex, ey, ez);
        double r2 = ex * ex + ey * ey + ez * ez, rd = sqrt(r2), t1, t2;
        rpy_real<double>(rd, R.alpha, t1, t2);
        double vx = vv[3 * j] * sc, vy = vv[3 * j + 1] * sc, vz = vv[3 * j + 2] * sc;
        double dv = (ex * vx + ey * vy + ez * vz) * t2 / r2;
        acc[0].add((T)(t1 * vx + dv * ex));
        acc[1].add((T)(t1 * vy + dv * ey));
        acc[2].add((T)(t1 * vz + dv * ez));
      }
    }
  }
  if constexpr (sizeof(T) == 8)
  for (int k = lane; k < m; k += 64) {
    int j = a.lrpy[(size_t)s * a.cap_rpy + k];
    double dx = a.rx[j] - xi, dy = a.ry[j] - yi, dz = a.rz[j] - zi;
    min_image_reduced(R, dx, dy, dz);
    T tx = (T)dx, ty = (T)dy, tz = (T)dz;
    T rr = tx * tx + ty * ty + tz * tz;
    if (rr < (T)4) {
      // overlapping pair: the overlap correction cancels 1/r^3 terms, so
      // evaluate it in FP64 whatever T is (rare in a physical configuration)
      double r2 = dx * dx + dy * dy + dz * dz, rd = sqrt(r2), t1, t2;
      rpy_real<double>(rd, R.alpha, t1, t2);
      double vx = vv[3 * j] * sc, vy = vv[3 * j + 1] * sc, vz = vv[3 * j + 2] * sc;
      double dv = (dx * vx + dy * vy + dz * vz) * t2 / r2;
      acc[0].add((T)(t1 * vx + dv * dx));
      acc[1].add((T)(t1 * vy + dv * dy));
      acc[2].add((T)(t1 * vz + dv * dz));
    } else if (rr < rc2) {
      T rd = sqrt(rr), t1, t2;
      rpy_real<T>(rd, alpha, t1, t2);
      T vx = (T)(vv[3 * j] * sc), vy = (T)(vv[3 * j + 1] * sc), vz = (T)(vv[3 * j + 2] * sc);
      T dv = (tx * vx + ty * vy + tz * vz) * t2 / rr;
      acc[0].add(t1 * vx + dv * tx);
      acc[1].add(t1 * vy + dv * ty);
      acc[2].add(t1 * vz + dv * tz);
    }
  }
  double u0 = warp_sum(acc[0].value()), u1 = warp_sum(acc[1].value()), u2 = warp_sum(acc[2].value());
  if (lane == 0) {
    double vi0 = vv[3 * s] * sc, vi1 = vv[3 * s + 1] * sc, vi2 = vv[3 * s + 2] * sc;
    a.u[3 * s] = u0 + R.selfconst * vi0;
    a.u[3 * s + 1] = u1 + R.selfconst * vi1;
    a.u[3 * s + 2] = u2 + R.selfconst * vi2;
  }
}

// Mixed precision, one workgroup (CL_WAVES wavefronts) per bead cluster I (<= 8 beads):
// the interpolation of each bead (one wavefront, a lane per stencil column), then
// the real-space pairs by cluster pairs: lane 8*ii + jj evaluates bead ii of I
// against bead jj of the listed cluster J, the wavefronts taking J in turn.  The
// partner data of one J is 8 positions and 8 vectors shared by the wave, so list
// and gather traffic fall by ~the cluster size against a bead-pair list.
// Coefficients come from the replica's FP64-fitted table, which includes the RPY
// overlap correction down to RTAB_R0; closer pairs take the FP32 closed form in
// erf, so no FP64 path (and none of its registers) is needed here.  With TAB_LDS
// the workgroup first copies the table to LDS (while its interpolation loads are
// in flight): the per-lane lookups are random, and as LDS reads they no longer
// queue behind one another in the vector memory path.  Sums are compensated per
// lane and combined in FP64 in a fixed order.
#define CL_WAVES 4
// Cluster pairs in flight per wave (CL_UNROLL, a divisor of 64) and the
// occupancy the register budget is held to: two pairs at 6 waves per SIMD ran
// 10% faster than four pairs at the 4 waves their registers allowed.
#ifndef CL_UNROLL
#define CL_UNROLL 2
#endif
#ifndef CL_WAVES_PER_EU
#define CL_WAVES_PER_EU 6
#endif
typedef __attribute__((address_space(3))) const float4 lds_float4;
template <int P, bool TAB_LDS>
__global__ __launch_bounds__(64 * CL_WAVES) __attribute__((amdgpu_waves_per_eu(CL_WAVES_PER_EU))) void k_mob_cl(
    MobArgs a) {
  typedef float T;
  extern __shared__ float4 stab[];  // 2 * ntab_max (TAB_LDS)
  const int I = a.cl_begin + xcd_block(blockIdx.x, gridDim.x);
  const int first = a.cl_first[I], cnt = a.cl_cnt[I];
  const int r = a.slot_rep[first];
  const RepDev& R = a.reps[r];
  if (!R.active || (a.done && a.done[r])) return;
  __shared__ double part[CL_WAVES][8][3];
  __shared__ double uin[8][3];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double* vv = a.vout ? a.vout : a.v;
  const double sc = (a.beta && !a.vout) ? (a.beta[r] > 0.0 ? 1.0 / a.beta[r] : 0.0) : 1.0;
  if constexpr (TAB_LDS) {
    const float4* g = a.rtab + R.tab0;
    for (int q = threadIdx.x; q < 2 * R.ntab; q += 64 * CL_WAVES) stab[q] = g[q];
  }
  // reciprocal part, bead ii by wavefront ii % CL_WAVES
  for (int ii = w; ii < cnt; ii += CL_WAVES) {
    const int s = first + ii;
    CompSum<T> ai[3];
    interp_bead<T, P>(a, R, s, lane, ai);
    double x0 = warp_sum(ai[0].value()), x1 = warp_sum(ai[1].value()), x2 = warp_sum(ai[2].value());
    if (lane == 0) {
      uin[ii][0] = x0;
      uin[ii][1] = x1;
      uin[ii][2] = x2;
    }
  }
  if constexpr (TAB_LDS) __syncthreads();
  // real-space part by cluster pairs.  The wave takes its list entries in groups
  // of CL_UNROLL consecutive ones and, every 64 entries, its lanes load one
  // entry each and compute its image shift: the minimum image of the vector
  // between the two clusters' first beads.  The pair loop (wave-uniform: it
  // starts from readfirstlane) broadcasts an entry and its shift from the lane
  // holding them with readlane, so it issues no list loads.  Positions are cluster-consistent:
  // a bead pair's vector under the shift is at most s = (that vector's length plus
  // both cluster radii) long, and every other image of it at least 2h - s, h being
  // half the shortest lattice vector.  When s <= 2h - rc no other image lies within
  // the cut-off, and a vector shorter than rc <= h is its own minimum image, so the
  // shift is exact for every pair the sum takes and the lanes skip their own
  // minimum-image reduction; the entry's sign bit marks the cluster pairs beyond
  // that bound (first slots fit 27 bits).
  // Lanes past the partner cluster's end load the slots that follow it (zeros
  // past the array, through the buffer's range check) and select the table's
  // zero entry, as do pairs beyond rc and the bead with itself, so the sums take
  // no selects.
  const int ii = lane >> 3, jj = lane & 7;
  const bool ivalid = ii < cnt;
  const unsigned long long ivm = cnt >= 8 ? ~0ull : (1ull << (8 * cnt)) - 1ull;  // lanes with ii < cnt
  const int si = first + (ivalid ? ii : 0);
  const float4 pi = a.p32[si];
  const float4 pc = a.p32[first];  // first bead; .w = cluster radius
  const float rd0 = (float)R.rd[0], rd1 = (float)R.rd[1], rd2 = (float)R.rd[2], rd3 = (float)R.rd[3];
  const float rb0 = (float)R.rb[0], rb1 = (float)R.rb[1], rb2 = (float)R.rb[2], rb3 = (float)R.rb[3];
  const float Lz = (float)R.L, iLz = (float)R.invL, rcf = (float)R.rc2;
  const float alpha = (float)R.alpha;
  // the bound 2h - rc on s, less a margin for the FP32 arithmetic of the test
  const float smax = (float)(2.0 * R.halfw - sqrt(R.rc2)) * 0.999f - 1e-3f;
  CompSum<T> acc[3];
  const int m = a.nclp[I];
  const int* row = a.lcl + (size_t)I * a.capc;
  const float rtop = rtab_top(R.ntab);
  // a VGPR copy, so the per-pair select reads one scalar operand (its mask) and the
  // compiler does not rematerialise the copy in the loop
  float vtop = rtop;
  asm volatile("" : "+v"(vtop));
  const unsigned stab_k = (unsigned)(size_t)(const lds_float4*)stab - 32u * RTAB_BASE;
  const __amdgpu_buffer_rsrc_t rp32 = __builtin_amdgcn_make_buffer_rsrc((void*)a.p32, (short)0, a.nst * 16, 0x00020000);
  const __amdgpu_buffer_rsrc_t rv32 = __builtin_amdgcn_make_buffer_rsrc((void*)a.v32, (short)0, a.nst * 16, 0x00020000);
  const int wu = __builtin_amdgcn_readfirstlane(w);
  constexpr int G = CL_UNROLL;
  static_assert(64 % G == 0, "a wave's 64 precomputed shifts cover whole groups of CL_UNROLL entries");
  for (int c0 = wu * G; c0 < m; c0 += 64 * CL_WAVES) {
    // lane l: entry c0 + CL_WAVES * G * (l / G) + l % G of this wave's next 64,
    // kept in the lane's registers and read by the pair loop with readlane
    int e_lane = 0;  // past the row: no beads
    float sh_x = 0.f, sh_y = 0.f, sh_z = 0.f;
    {
      const int e_idx = c0 + CL_WAVES * G * (lane / G) + lane % G;
      if (e_idx < m) {
        e_lane = row[e_idx];
        const float4 pr = a.p32[e_lane >> 4];
        float cx = pr.x - pc.x, cy = pr.y - pc.y, cz = pr.z - pc.z;
        const float n1 = rintf(rd0 * cx + rd1 * cy), n2 = rintf(rd2 * cx + rd3 * cy), n3 = rintf(cz * iLz);
        const float sx = -(n1 * rb0 + n2 * rb2), sy = -(n1 * rb1 + n2 * rb3), sz = -Lz * n3;
        cx += sx;
        cy += sy;
        cz += sz;
        const float reach = __builtin_amdgcn_sqrtf(cx * cx + cy * cy + cz * cz) + pr.w + pc.w;
        sh_x = sx;
        sh_y = sy;
        sh_z = sz;
        if (!(reach < smax)) e_lane |= (int)0x80000000u;
      }
    }
    const int c1 = min(m, c0 + 64 * CL_WAVES);
    // the chunk's (<= 64) terms of each lane are summed plainly in FP32, then
    // added to the lane's compensated sum
    float px = 0.f, py = 0.f, pz = 0.f;
    for (int k0 = c0, q = 0; k0 < c1; k0 += CL_WAVES * G, q += G) {
      int sj[G];
      bool ok[G], zs[G];
      float3 pj[G], vj[G];
      float4 S[G];
      int cflag = 0;
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int e = __builtin_amdgcn_readlane(e_lane, q + u);  // 0 past the row's end
        const int fj = (e >> 4) & 0x07ffffff, cj = e & 15;
        // lanes of a real bead pair as a wave mask on the scalar unit: bead ii
        // of I present, bead jj < cj of J, not the bead with itself
        unsigned long long om = ivm & ((0x0101010101010101ull << cj) - 0x0101010101010101ull);
        if (fj == first) om &= ~0x8040201008040201ull;
        ok[u] = __builtin_amdgcn_inverse_ballot_w64(om);
        sj[u] = fj + jj;
        const int ix = __builtin_amdgcn_readlane(__builtin_bit_cast(int, sh_x), q + u);
        const int iy = __builtin_amdgcn_readlane(__builtin_bit_cast(int, sh_y), q + u);
        const int iz = __builtin_amdgcn_readlane(__builtin_bit_cast(int, sh_z), q + u);
        S[u] = make_float4(__builtin_bit_cast(float, ix), __builtin_bit_cast(float, iy), __builtin_bit_cast(float, iz), 0.f);
        zs[u] = ((ix | iy | iz) & 0x7fffffff) == 0;  // on the scalar unit: +-0 in every component
        cflag |= e;
      }
      const bool compact = cflag >= 0;
#pragma unroll
      for (int u = 0; u < G; ++u) {  // 32-bit offsets into the position and vector copies
        const auto lp = __builtin_amdgcn_raw_buffer_load_b96(rp32, sj[u] * 16, 0, 0);
        const auto lv = __builtin_amdgcn_raw_buffer_load_b96(rv32, sj[u] * 16, 0, 0);
        pj[u] = make_float3(__int_as_float(lp[0]), __int_as_float(lp[1]), __int_as_float(lp[2]));
        vj[u] = make_float3(__int_as_float(lv[0]), __int_as_float(lv[1]), __int_as_float(lv[2]));
      }
#pragma unroll
      for (int u = 0; u < G; ++u) {
        float dx, dy, dz;
        if (compact && zs[u]) {  // the clusters' common image is the unshifted one
          dx = pj[u].x - pi.x;
          dy = pj[u].y - pi.y;
          dz = pj[u].z - pi.z;
        } else if (compact) {
          dx = (pj[u].x + S[u].x) - pi.x;
          dy = (pj[u].y + S[u].y) - pi.y;
          dz = (pj[u].z + S[u].z) - pi.z;
        } else {
          dx = pj[u].x - pi.x;
          dy = pj[u].y - pi.y;
          dz = pj[u].z - pi.z;
          const float n1 = rintf(rd0 * dx + rd1 * dy), n2 = rintf(rd2 * dx + rd3 * dy);
          dx -= n1 * rb0 + n2 * rb2;
          dy -= n1 * rb1 + n2 * rb3;
          dz -= Lz * rintf(dz * iLz);
        }
        const float rr = dx * dx + dy * dy + dz * dz;
        // Table coefficients on every lane: r^2 clamped into [1, rc^2] (the cut-off
        // is a table knot, where the zero entry starts; pairs not summed clamp to
        // it from below), the entry read from the bits, so the common path has no
        // branch and the partner loads stay ahead of it; pairs closer than the
        // table take the closed form (rare).
        const float rs = __builtin_amdgcn_fmed3f(rr, ok[u] ? 1.f : vtop, rtop);
        const unsigned b = __float_as_uint(rs);
        const float t = __uint_as_float(((b << RTAB_MBITS) & 0x007fffffu) | 0x3f800000u);
        float4 ca, cb;
        if constexpr (TAB_LDS) {
          // the entry's LDS address from the unreduced segment bits (stab_k is
          // the table's address less RTAB_BASE entries)
          unsigned addr;  // (b >> RTAB_SHIFT) * 32 + stab_k in one instruction
          asm("v_lshl_add_u32 %0, %1, 5, %2" : "=v"(addr) : "v"(b >> RTAB_SHIFT), "s"(stab_k));
          const float4* e = (const float4*)(const lds_float4*)(size_t)addr;
          ca = e[0];
          cb = e[1];
        } else {
          const int it = (int)((b >> RTAB_SHIFT) - RTAB_BASE);
          ca = a.rtab[R.tab0 + 2 * it];
          cb = a.rtab[R.tab0 + 2 * it + 1];
        }
        float t1 = ca.x + t * (ca.y + t * (ca.z + t * ca.w));
        float b2 = cb.x + t * (cb.y + t * (cb.z + t * cb.w));
        if (ok[u] && rr < RTAB_R0 * RTAB_R0) {
          float t2;
          rpy_real_near_f32(sqrtf(rr), rr, alpha, t1, t2);
          b2 = t2 / rr;
        }
        const float dv = (dx * vj[u].x + dy * vj[u].y + dz * vj[u].z) * b2;
        px = fmaf(dv, dx, fmaf(t1, vj[u].x, px));
        py = fmaf(dv, dy, fmaf(t1, vj[u].y, py));
        pz = fmaf(dv, dz, fmaf(t1, vj[u].z, pz));
      }
    }
    acc[0].add(px);
    acc[1].add(py);
    acc[2].add(pz);
  }
  double u0 = acc[0].value(), u1 = acc[1].value(), u2 = acc[2].value();
  for (int o = 1; o < 8; o <<= 1) {
    u0 += __shfl_xor(u0, o, 64);
    u1 += __shfl_xor(u1, o, 64);
    u2 += __shfl_xor(u2, o, 64);
  }
  if (jj == 0) {
    part[w][ii][0] = u0;
    part[w][ii][1] = u1;
    part[w][ii][2] = u2;
  }
  __syncthreads();
  double pr = 0.0;  // this thread's term of v.u
  if (threadIdx.x < 3 * cnt) {
    const int b = threadIdx.x / 3, c = threadIdx.x - 3 * b, s = first + b;
    double t = uin[b][c];
    for (int q = 0; q < CL_WAVES; ++q) t += part[q][b][c];
    const double vi = vv[3 * s + c] * sc;
    t += R.selfconst * vi;
    a.u[3 * s + c] = t;
    pr = vi * t;
  }
  if (a.adot && threadIdx.x < 64) {  // the cluster's share of the Lanczos alpha, (bead, component) order
    pr = warp_sum(pr);
    if (threadIdx.x == 0) a.adot[I] = pr;
  }
}

template <typename T, int P> static void launch_mob_p(hipStream_t st, const MobArgs& a) {
  auto mark = [&](int k) {
    if (a.ev) (void)hipEventRecord(a.ev[k], st);
  };
  mark(0);
  int nt = (a.nslots + 255) / 256;
  // The spread normalises the input itself when it runs in this launch and the
  // range is small: that saves a launch per product where launches dominate
  // (-2.5% on an 8000-bead step), while over 32k slots the staged FP64 reads
  // and stores cost more than the launch (+2.6% on a 380k-bead batch).
  const bool fuse = (a.phase & 3) == 3 && a.nbricks > 0 && (a.vout || a.v32) && a.nslots <= SCALE_FUSE_MAX_SLOTS;
  if (a.phase & 1) {  // per-bead preparation over the whole replica range
    if (a.prep && a.nslots > 0) {
      launch_fill_i32(st, a.brick_cnt + a.brick_begin, 0, a.brick_len);
      launch_fill_i32(st, a.brick_fill + a.brick_begin, 0, a.brick_len);
      k_bead_prep<T, P><<<nt, 256, 0, st>>>(a);
      if (a.p32 && a.lcl && a.ncl > 0) k_cluster_prep<<<(a.ncl + 255) / 256, 256, 0, st>>>(a);
      launch_scan(st, a.brick_cnt + a.brick_begin, a.brick_start + a.brick_begin, a.brick_len);
      k_brick_fill<<<nt, 256, 0, st>>>(a);
      if (a.brick_len > 0) k_brick_sort<<<(a.brick_len + 255) / 256, 256, 0, st>>>(a);
      if (a.bcand && a.brick_len > 0) {
        k_brick_cand_count<<<(a.brick_len + 255) / 256, 256, 0, st>>>(a);
        launch_scan(st, a.bcand_cnt + a.brick_begin, a.bcand_start + a.brick_begin, a.brick_len);
        k_brick_cand_fill<<<a.brick_len, 64, 0, st>>>(a);
      }
    }
    if ((a.vout || a.v32) && a.nslots > 0 && !fuse) k_scale_v<<<nt, 256, 0, st>>>(a);
  }
  mark(1);
  if (!(a.phase & 2)) return;
  if (a.nbricks > 0) {
    MobArgs as = a;
    as.scale_in_spread = fuse ? 1 : 0;
    if (sizeof(T) == 4 && a.bcand)
      k_spread_mfma<P><<<a.nbricks, 256, 0, st>>>(as);
    else
      k_spread_gather<T, P><<<a.nbricks, 512, 0, st>>>(as);
  }
  mark(2);
  // The grid passes, chunk by chunk over the replicas (one chunk: the whole range).
  // One chunk is timed by the marks 3-6 between its passes; several chunks each
  // record six events of their own around their five passes (evc), and the marks
  // 3-6 follow the last chunk.
  const int nch = a.nfch > 1 ? a.nfch : 1;
  for (int q = 0; q < nch; ++q) {
    auto pmark = [&](int k) {
      if (nch == 1) {
        if (k >= 1 && k <= 4) mark(2 + k);
      } else if (a.evc) {
        (void)hipEventRecord(a.evc[6 * q + k], st);
      }
    };
    MobArgs f = a;
    if (nch > 1) {
      const int *b = a.fch + 4 * q, *e = b + 4;
      f.tx = a.tx + b[0];
      f.ty = a.ty + b[1];
      f.tz = a.tz + b[2];
      f.txy = a.txy + b[3];
      f.ntx = e[0] - b[0];
      f.nty = e[1] - b[1];
      f.ntz = e[2] - b[2];
      f.ntxy = e[3] - b[3];
    }
    pmark(0);
    if (f.ntxy > 0) k_fft_xy<T><<<f.ntxy, XY_THREADS, 0, st>>>(f, -1);
    if (f.ntx > 0) k_fft_x<T><<<f.ntx, FFT_THREADS, 0, st>>>(f, -1);
    pmark(1);
    if (f.nty > 0) k_fft_y<T><<<f.nty, FFT_THREADS, 0, st>>>(f, -1);
    pmark(2);
    if (f.ntz > 0 && f.ztile == FFT_TILE_Z_LARGE)
      k_fft_z_green<T, FFT_TILE_Z_LARGE><<<f.ntz, FFT_THREADS, 0, st>>>(f);
    else if (f.ntz > 0 && f.ztile == FFT_TILE_Z)
      k_fft_z_green<T, FFT_TILE_Z><<<f.ntz, FFT_THREADS, 0, st>>>(f);
    else if (f.ntz > 0)
      k_fft_z_green<T, FFT_TILE_Z_SMALL><<<f.ntz, FFT_THREADS, 0, st>>>(f);
    pmark(3);
    if (f.nty > 0) k_fft_y<T><<<f.nty, FFT_THREADS, 0, st>>>(f, 1);
    pmark(4);
    if (f.ntxy > 0) k_fft_xy<T><<<f.ntxy, XY_THREADS, 0, st>>>(f, 1);
    if (f.ntx > 0) k_fft_x<T><<<f.ntx, FFT_THREADS, 0, st>>>(f, 1);
    pmark(5);
  }
  if (nch > 1) {
    mark(3);
    mark(4);
    mark(5);
    mark(6);
  }
  mark(7);
  if (sizeof(T) == 4 && a.lcl) {
    if (a.ncl > 0 && a.ntab_max <= RTAB_LDS_MAX)
      k_mob_cl<P, true><<<a.ncl, 64 * CL_WAVES, 32 * a.ntab_max, st>>>(a);
    else if (a.ncl > 0)
      k_mob_cl<P, false><<<a.ncl, 64 * CL_WAVES, 0, st>>>(a);
  } else if (a.nslots > 0) {
    k_mob_final<T, P><<<(a.nslots + 3) / 4, 256, 0, st>>>(a);
  }
  mark(8);
}

template <typename T> void launch_mobility(hipStream_t st, const MobArgs& a) {
  if (a.order == 4)
    launch_mob_p<T, 4>(st, a);
  else if (a.order == 8)
    launch_mob_p<T, 8>(st, a);
  else
    launch_mob_p<T, 6>(st, a);
}

template void launch_mobility<float>(hipStream_t, const MobArgs&);
template void launch_mobility<double>(hipStream_t, const MobArgs&);
