// Adaptive widths per phi block on the device (xpg_adapt_block, include/xpg_integrate.h): the small fits of
// xpg/integrate.py:404-461 as a chain of six launches on the caller's stream, so the per-block chain
// pass 1 -> fits -> pass 2 -> next block runs without a host round trip.
//
//   k_ad_prep    (n/256 x 256)   strong-row flags, fz totals and sigma_D terms per row; strong rows per workgroup
//   k_ad_grid    (NGW x 256)     partial cost of each sigma_M of the 61-point log grid over fixed ranges of strong
//                                rows, one wavefront per row and one lane per grid point (_fit_rocking_width,
//                                :1130-1154)
//   k_ad_finish  (1 x 256)       strong-row count, argmin over the grid -> sigma_M fit and shrinkage
//   k_ad_gram    (NG x 64)       Gram matrix A^T A and A^T y of the rocking-curve basis over a fixed row range
//                                (rc_basis_np, :437-443)
//   k_ad_gsum    (NE + 1 x 256)  fixed-order reduction of the Gram partials, one workgroup per entry; the last one
//                                selects the exact median of the sigma_D terms (radix selection of the two middle
//                                ranks among the strong keys k_ad_grid laid out)
//   k_ad_solve   (1 x 64)        non-negative least squares on the Gram form (_nnls_gram, :1095-1104) in one
//                                wavefront, regularisation towards the Gaussian weights, the widths records
//
// Every row range and every reduction order is fixed by the row count alone: results do not depend on scheduling.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "xpg_erf.h"
#include "xpg_integrate.h"

namespace {

constexpr int K = XPG_RC_K;               // 21 mixture components
constexpr int NF = 65;                     // frames of a per-frame profile row (max(DZ_BUCKETS))
constexpr int NG = 4096;                   // workgroups of the Gram accumulation (one wavefront each)
constexpr int NGW = 512;                   // workgroups of the grid costs (4 wavefronts, one strong-row range each)
constexpr int NE = K * (K + 1) / 2 + K;    // unique Gram entries + right-hand side
constexpr int NGRID = 61;
static_assert(NG % 256 == 0 && NGW % 64 == 0, "the fixed-order reductions read whole chunks");
constexpr double PI = 3.141592653589793;

// rc_default_weights (float32 values of exp(-c^2 / (2 (1 - 0.25^2))) normalised)
__device__ __constant__ static const double RC_C[K] = {
    -3.5, -3.15, -2.8, -2.4499999999999997, -2.1, -1.75, -1.4, -1.0499999999999998, -0.7, -0.34999999999999964, 0.0,
    0.35000000000000053, 0.7000000000000002, 1.0500000000000007, 1.4000000000000004, 1.75, 2.1000000000000005,
    2.4500000000000006, 2.8000000000000007, 3.1500000000000004, 3.5};

struct Ws {
  uint8_t *st;        // n: strong-row flag
  uint64_t *key;      // n: order-preserving bits of the sigma_D term
  double *tot;        // n: fz total
  int32_t *nzl;       // n: 1 + the last frame with fz != 0 (0: none)
  int32_t *cnt;       // nprep: strong rows per k_ad_prep workgroup
  double *cost;       // NGW x 64 partial grid costs (lane c < NGRID: grid point c)
  double *gram;       // NG x NE partial Gram entries
  double *gtot;       // NE Gram totals
  double *scal;       // n_strong, sm_new, sm_fit, sd_new, sd_fit
  uint64_t *kmm;      // nprep x 2: smallest and largest strong key per k_ad_prep workgroup
  uint64_t *kc;       // n: the strong keys in row order (k_ad_grid)
  int32_t nprep;
};

__host__ __device__ inline Ws carve(void *base, int32_t n) {
  Ws w;
  char *p = (char *)base;
  auto take = [&](size_t bytes) {
    char *q = p;
    p += (bytes + 255) & ~(size_t)255;
    return q;
  };
  const size_t m = (size_t)(n > 0 ? n : 1);
  w.nprep = (int32_t)((m + 255) / 256);
  w.st = (uint8_t *)take(m);
  w.key = (uint64_t *)take(8 * m);
  w.tot = (double *)take(8 * m);
  w.nzl = (int32_t *)take(4 * m);
  w.cnt = (int32_t *)take(4 * (size_t)w.nprep);
  w.cost = (double *)take(8 * NGW * 64);
  w.gram = (double *)take(8 * (size_t)NG * NE);
  w.gtot = (double *)take(8 * (size_t)NE);
  w.scal = (double *)take(8 * 8);
  w.kmm = (uint64_t *)take(16 * (size_t)w.nprep);
  w.kc = (uint64_t *)take(8 * m);
  return w;
}

struct In {
  int32_t n;
  const double *isum, *vsum, *frac_rec, *m_d, *phi, *zeta_k, *Xlen, *cosa;
  const int32_t *n_over, *iz0, *fb;
  const float *fz;
  double sd_run, sm_run, phi0, osc, qx;
};

__device__ __forceinline__ double bound_deg(const In &a, int iz0, int k) { return a.phi0 + a.osc * (double)(iz0 + k); }

// rows [r0, r1) of range c of nc equal ranges
__device__ __forceinline__ void row_range(int n, int c, int nc, int &r0, int &r1) {
  const int per = (n + nc - 1) / nc;
  r0 = min(c * per, n);
  r1 = min(r0 + per, n);
}

__device__ __forceinline__ double rld(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// fixed-shape tree sum over a workgroup of 256 threads
__device__ __forceinline__ double block_sum256(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// ---- 1. strong rows -------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ad_prep(In a, Ws w) {
  __shared__ int32_t wc[4];
  __shared__ uint64_t wm[4][2];
  const int r = blockIdx.x * 256 + threadIdx.x;
  bool st = false;
  uint64_t kk = 0;
  if (r < a.n) {
    st = (a.isum[r] > 0) && (a.isum[r] / sqrt(fmax(a.vsum[r], 1e-6)) > 15.0) && (a.n_over[r] == 0) &&
         (a.frac_rec[r] > 0.98) && (a.fb[r] == 1);
    double s = 0.0, md = 0.0;
    int last = 0;
    if (st) {
      for (int k = 0; k < NF; ++k) {
        const float f = a.fz[(int64_t)r * NF + k];
        s += (double)f;
        if (f != 0.f) last = k + 1;
      }
      const double t2 = 1.0 - (9.0 / 2) * exp(-9.0 / 2) / (1 - exp(-9.0 / 2));
      const double q = a.qx / a.Xlen[r];
      md = a.m_d[r] / t2 - (q * q) * (1 + a.cosa[r] * a.cosa[r]) / 24.0;
    }
    const uint64_t b = __builtin_bit_cast(uint64_t, md);
    kk = (b >> 63) ? ~b : (b | 0x8000000000000000ull);            // unsigned order = numeric order
    w.key[r] = kk;
    w.tot[r] = s;
    w.nzl[r] = last;
    w.st[r] = st;
  }
  const uint64_t bal = __ballot(st);
  uint64_t kmin = ~0ull, kmax = 0;
  if (st) kmin = kmax = kk;
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a0 = __shfl_xor(kmin, o, 64), a1 = __shfl_xor(kmax, o, 64);
    kmin = a0 < kmin ? a0 : kmin;
    kmax = a1 > kmax ? a1 : kmax;
  }
  if ((threadIdx.x & 63) == 0) {
    wc[threadIdx.x >> 6] = __popcll(bal);
    wm[threadIdx.x >> 6][0] = kmin;
    wm[threadIdx.x >> 6][1] = kmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    w.cnt[blockIdx.x] = (wc[0] + wc[1]) + (wc[2] + wc[3]);
    uint64_t lo = wm[0][0], hi = wm[0][1];
    for (int i = 1; i < 4; ++i) {
      lo = wm[i][0] < lo ? wm[i][0] : lo;
      hi = wm[i][1] > hi ? wm[i][1] : hi;
    }
    w.kmm[2 * blockIdx.x] = lo;
    w.kmm[2 * blockIdx.x + 1] = hi;
  }
}

// The sigma_D median runs in pieces across the chain, in small workgroups: k_ad_prep gives the strong keys' range,
// k_ad_grid lays the strong keys out in row order, and one workgroup of k_ad_gsum selects the two middle ranks.

// smallest and largest strong key (one wavefront; every lane returns them), the byte of their highest differing bit
// and their common prefix above it
__device__ __forceinline__ void key_range(const Ws &w, uint64_t &kmin, uint64_t &kmax, int &top, uint64_t &pre) {
  const int lane = threadIdx.x & 63;
  kmin = ~0ull;
  kmax = 0;
  for (int b = lane; b < w.nprep; b += 64) {
    const uint64_t lo = w.kmm[2 * b], hi = w.kmm[2 * b + 1];
    kmin = lo < kmin ? lo : kmin;
    kmax = hi > kmax ? hi : kmax;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a0 = __shfl_xor(kmin, o, 64), a1 = __shfl_xor(kmax, o, 64);
    kmin = a0 < kmin ? a0 : kmin;
    kmax = a1 > kmax ? a1 : kmax;
  }
  const uint64_t diff = kmin ^ kmax;
  top = diff ? (63 - __clzll((long long)diff)) / 8 * 8 : 0;
  pre = top == 56 ? 0ull : (kmin & (~0ull << (top + 8)));
}

// ---- 2. grid of sigma_M candidates -------------------------------------------------------------------------------

// erf(x); beyond |x| = ERF_SAT it is +-1 exactly (float64 erf rounds to +-1 there, and erf_tab returns it)
constexpr double ERF_SAT = 6.0;
__device__ __forceinline__ double erf_sat(double x) { return xpg::erf_tab(x); }

// ---- 3. sigma_M from the grid, sigma_D from the median ----------------------------------------------------------

// the bin of a 256-bin histogram holding rank k (one wavefront): returns the bin, k reduced to the rank inside it
__device__ __forceinline__ int select_bin(const uint32_t *h, uint32_t &k) {
  const int lane = threadIdx.x & 63;
  const uint32_t c0 = h[4 * lane], c1 = h[4 * lane + 1], c2 = h[4 * lane + 2], c3 = h[4 * lane + 3];
  const uint32_t s = (c0 + c1) + (c2 + c3);
  uint32_t incl = s;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const uint64_t hit = __ballot(incl > k);
  const int L = __ffsll((long long)hit) - 1;        // the rank exists: L >= 0
  uint32_t excl = __shfl(incl - s, L, 64);
  const uint32_t cc[4] = {__shfl(c0, L, 64), __shfl(c1, L, 64), __shfl(c2, L, 64), __shfl(c3, L, 64)};
  int bin = 4 * L + 3;
  for (int i = 0; i < 4; ++i) {
    if (excl + cc[i] > k) {
      bin = 4 * L + i;
      break;
    }
    excl += cc[i];
  }
  k -= excl;
  return bin;
}

// histogram increment of a wavefront (all 64 lanes call it): one atomic when every live lane hits the same bin
__device__ __forceinline__ void hist_add(uint32_t *h, uint32_t bin, bool live) {
  const uint64_t m = __ballot(live);
  if (!m) return;
  const int first = __ffsll((long long)m) - 1;
  const uint32_t lead = __shfl(bin, first, 64);
  const uint64_t same = __ballot(live && bin == lead);
  if (same == m) {
    if ((int)(threadIdx.x & 63) == first) atomicAdd(&h[lead], (uint32_t)__popcll(m));
  } else if (live) {
    atomicAdd(&h[bin], 1u);
  }
}

// Exact median of the strong rows' sigma_D terms in one workgroup of MT = 256 threads: radix selection, 8 bits per
// round from the first byte where the keys differ, of the two middle ranks among the n strong keys laid out by
// k_ad_grid (held in registers, KR per thread, when n <= MT KR); once at most 64 keys share a rank's prefix, that rank
// is finished among them directly.  sigma_D fit and shrinkage into scal[3], scal[4].
__device__ void median_select(const In &a, const Ws &w) {
  constexpr int MT = 256, KR = 24;
  __shared__ uint32_t hist[2][256];
  __shared__ uint64_t s_pre[2], s_msk[2];  // rank q: the bits of its key fixed so far
  __shared__ uint32_t s_k[2];              // rank q: its rank among the keys with that prefix
  __shared__ int32_t s_done[2], s_cn[2];
  __shared__ uint64_t s_cand[2][64];
  const int t = threadIdx.x;
  const int n = (int)w.scal[0];
  const bool two = !(n & 1);               // the two middle ranks differ
  uint64_t kmin, kmax, pre;
  int top;
  key_range(w, kmin, kmax, top, pre);
  if (t == 0) {
    s_pre[0] = s_pre[1] = pre;
    s_msk[0] = s_msk[1] = top == 56 ? 0ull : (~0ull << (top + 8));
    s_k[0] = (uint32_t)((n - 1) / 2);
    s_k[1] = (uint32_t)(n / 2);
    s_done[0] = s_done[1] = 0;
    s_cn[0] = s_cn[1] = 0;
  }
  const bool in_regs = n <= MT * KR;
  uint64_t kr[KR];
#pragma unroll
  for (int i = 0; i < KR; ++i) {
    const int j = t + MT * i;
    kr[i] = in_regs && j < n ? w.kc[j] : 0ull;
  }
  __syncthreads();
  // one pass over the keys, all 64 lanes of every wavefront together (f takes a liveness flag)
  auto for_keys = [&](auto &&f) {
    if (in_regs) {
#pragma unroll
      for (int i = 0; i < KR; ++i) f(kr[i], t + MT * i < n);
    } else {
      for (int i0 = 0; i0 < n; i0 += MT) {
        const int i = i0 + t;
        const bool live = i < n;
        f(live ? w.kc[i] : 0ull, live);
      }
    }
  };
  for (int shift = top; shift >= 0 && !(s_done[0] && s_done[1]); shift -= 8) {
    for (int i = t; i < 512; i += MT) (&hist[0][0])[i] = 0;
    __syncthreads();
    const uint64_t hm = shift == 56 ? 0ull : (~0ull << (shift + 8));
    const uint64_t p0 = s_pre[0], p1 = s_pre[1];
    const bool a0 = !s_done[0], a1 = two && !s_done[1];
    for_keys([&](uint64_t key, bool live) {
      const uint32_t bin = (uint32_t)(key >> shift) & 255u;
      if (a0) hist_add(hist[0], bin, live && (key & hm) == p0);
      if (a1) hist_add(hist[1], bin, live && (key & hm) == p1);
    });
    __syncthreads();
    if (t < (two ? 128 : 64) && !s_done[t >> 6]) {
      const int q = t >> 6;
      uint32_t k = s_k[q];
      const int bin = select_bin(hist[q], k);
      if ((t & 63) == 0) {
        s_pre[q] |= (uint64_t)bin << shift;
        s_msk[q] = ~0ull << shift;
        s_k[q] = k;
        s_done[q] = (int32_t)hist[q][bin] <= 64;
        if (!two) s_done[1] = 1;
      }
    }
    __syncthreads();
  }
  // ranks with bits still open: gather the keys with the rank's prefix (at most 64; any order), then the one whose
  // rank among them (ties broken by slot) is the remaining rank
  const bool g0 = s_msk[0] != ~0ull, g1 = two && s_msk[1] != ~0ull;
  if (g0 || g1) {
    const uint64_t m0 = s_msk[0], m1 = s_msk[1], p0 = s_pre[0], p1 = s_pre[1];
    for_keys([&](uint64_t key, bool live) {
      if (live && g0 && (key & m0) == p0) {
        const int i = atomicAdd(&s_cn[0], 1);
        if (i < 64) s_cand[0][i] = key;
      }
      if (live && g1 && (key & m1) == p1) {
        const int i = atomicAdd(&s_cn[1], 1);
        if (i < 64) s_cand[1][i] = key;
      }
    });
    __syncthreads();
    if (t < 128) {
      const int q = t >> 6, l = t & 63;
      if (q == 0 ? g0 : g1) {
        const int cn2 = min(s_cn[q], 64);
        const uint64_t c = l < cn2 ? s_cand[q][l] : ~0ull;
        int rank = 0;
        for (int j = 0; j < cn2; ++j) {
          const uint64_t o = s_cand[q][j];
          rank += (o < c) || (o == c && j < l);
        }
        if (l < cn2 && rank == (int)s_k[q]) s_pre[q] = c;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    double v[2];
    for (int q = 0; q < 2; ++q) {
      const uint64_t b = s_pre[two ? q : 0];
      v[q] = __builtin_bit_cast(double, (b >> 63) ? (b & 0x7fffffffffffffffull) : ~b);
    }
    const double med = (n & 1) ? v[1] : (v[0] + v[1]) / 2.0;
    double sd_fit = sqrt(fmax(med, 1e-12));
    sd_fit = fmin(fmax(sd_fit, 0.5 * a.sd_run), 2.0 * a.sd_run);
    w.scal[3] = exp(((double)n * log(sd_fit) + 300.0 * log(a.sd_run)) / ((double)n + 300.0));
    w.scal[4] = sd_fit;
  }
}

// strong-row count (scal[0]) and the sigma_M fit: the grid costs summed over the workgroups in order, the first
// minimum, shrinkage towards the running value (scal[1], scal[2])
__global__ void __launch_bounds__(256) k_ad_finish(In a, Ws w) {
  __shared__ int32_t s_n;
  __shared__ double s_part[4][64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t == 0) s_n = 0;
  __syncthreads();
  int cn = 0;
  for (int b = t; b < w.nprep; b += 256) cn += w.cnt[b];
  if (cn) atomicAdd(&s_n, cn);           // integer counts: order-free
  // wavefront q sums a quarter of the workgroup partials in order, sixteen reads in flight
  double part = 0.0;
  for (int g0 = wv * (NGW / 4); g0 < (wv + 1) * (NGW / 4); g0 += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = w.cost[(g0 + u) * 64 + lane];
#pragma unroll
    for (int u = 0; u < 16; ++u) part += v[u];
  }
  s_part[wv][lane] = part;
  __syncthreads();
  const int n = s_n;
  if (t == 0) w.scal[0] = (double)n;
  if (n < 100 || wv > 0) return;
  const double cst = ((s_part[0][lane] + s_part[1][lane]) + s_part[2][lane]) + s_part[3][lane];
  // first minimum over the grid points (lanes 0 .. NGRID - 1) below +inf; none: the running value
  double best = lane < NGRID ? cst : INFINITY;
  for (int o = 32; o > 0; o >>= 1) best = fmin(best, __shfl_xor(best, o, 64));
  const uint64_t hit = __ballot(lane < NGRID && cst == best && cst < INFINITY);
  const int c = hit ? __ffsll((long long)hit) - 1 : -1;
  if (lane == 0) {
    const double guess = a.sm_run * (180.0 / PI);
    const double lo_l = log(0.5), hi_l = log(2.0), step = (hi_l - lo_l) / (NGRID - 1);
    const double bsm = c < 0 ? guess : guess * exp(c == NGRID - 1 ? hi_l : (double)c * step + lo_l);
    const double sm_fit = bsm * (PI / 180.0);
    w.scal[1] = exp(((double)n * log(sm_fit) + 300.0 * log(a.sm_run)) / ((double)n + 300.0));
    w.scal[2] = sm_fit;
  }
}

// Grid costs: one wavefront per strong row (a fixed range of strong ranks per wavefront, in order), lane c for grid
// point c.
// The row's frame totals are one vector load (lane k: frame k < 64) broadcast by readlane; the row fields are uniform.
__global__ void __launch_bounds__(256) k_ad_grid(In a, Ws w) {
  __shared__ double wpart[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = blockIdx.x;
  const int c = lane < NGRID ? lane : NGRID - 1;      // lanes past the grid repeat its last point (not summed)
  const double guess = a.sm_run * (180.0 / PI);
  const double lo_l = log(0.5), hi_l = log(2.0), step = (hi_l - lo_l) / (NGRID - 1);
  const double smv = guess * exp(c == NGRID - 1 ? hi_l : (double)c * step + lo_l);
  // The strong rows in row order; this wavefront takes strong ranks [q0, q1).  The k_ad_prep workgroup counts locate
  // the first; windows of 64 rows (ballots of the strong flags) step through the rest.
  int ns = 0;
  for (int b0 = 0; b0 < w.nprep; b0 += 64) {
    int cv = b0 + lane < w.nprep ? w.cnt[b0 + lane] : 0;
    for (int o = 32; o > 0; o >>= 1) cv += __shfl_xor(cv, o, 64);
    ns += cv;
  }
  int q0, q1;
  row_range(ns, g * 4 + wv, NGW * 4, q0, q1);
  int need = q1 - q0, skip = q0, base = 0;
  for (int b0 = 0; need > 0 && b0 < w.nprep; b0 += 64) {
    const int cv = b0 + lane < w.nprep ? w.cnt[b0 + lane] : 0;
    int incl = cv;
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    const int all = __shfl(incl, 63, 64);
    if (skip < all) {
      const int L = __ffsll((long long)__ballot(incl > skip)) - 1;
      skip -= __shfl(incl - cv, L, 64);
      base = (b0 + L) * 256;
      break;
    }
    skip -= all;
  }
  int rank = q0;                         // the strong rank of the next row taken
  uint64_t m = 0;
  base -= 64;
  double part = 0.0;
  while (need > 0) {
    if (!m) {
      base += 64;
      if (base >= a.n) break;            // not reached: the counts are the flags' popcounts
      m = __ballot(base + lane < a.n && w.st[base + lane]);
      continue;
    }
    const int r = base + __ffsll((long long)m) - 1;
    m &= m - 1;
    if (skip > 0) {
      --skip;
      continue;
    }
    --need;
    // r is uniform, and so are the row's fields
    if (lane == 0) w.kc[rank] = w.key[r];  // the median's keys, in row order
    ++rank;
    const double tot = w.tot[r], phi = a.phi[r], zeta = fabs(a.zeta_k[r]);
    const int iz0 = a.iz0[r], kend = w.nzl[r];
    const float *fz = a.fz + (int64_t)r * NF;
    const float fzl = fz[lane];          // frames 0 .. 63; frame 64 below
    const float fz64 = fz[NF - 1];
    if (!(tot > 0)) continue;
    const double s = smv / zeta * sqrt(2.0);
    const double inv_s = 1.0 / s, inv_tot = 1.0 / tot;
    // frame k spans [lo, hi] = sorted(phi0 + osc (iz0 + k), phi0 + osc (iz0 + k + 1)); the row total of R telescopes
    const double e0 = erf_sat((bound_deg(a, iz0, 0) - phi) * inv_s), e1 = erf_sat((bound_deg(a, iz0, NF) - phi) * inv_s);
    const double den = fmax(a.osc >= 0 ? 0.5 * (e1 - e0) : 0.5 * (e0 - e1), 1e-9), inv_den = 1.0 / den;
    // Frames in chunks of 8: the chunk's boundary erfs are evaluated together (independent table reads), then its
    // terms are added in frame order.  The boundaries move monotonically away from phi: once boundary k0 is past the
    // saturation point in that direction every later erf is the same +-1, so with k0 past the last non-zero fz every
    // later term is tot * (+0)^2 and the lane's cost is final; the chunk loop stops when that holds for every lane.
    double cst = 0.0, Eprev = e0;
    for (int k0 = 0; k0 < NF; k0 += 8) {
      const double xb = (bound_deg(a, iz0, k0) - phi) * inv_s;
      const bool done = k0 >= kend && fabs(xb) > ERF_SAT && (xb > 0) == (a.osc >= 0);
      if (!__ballot(!done)) break;
      double E[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) E[u] = erf_sat((bound_deg(a, iz0, min(k0 + u + 1, NF)) - phi) * inv_s);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        if (k < NF) {
          const float fk = k < NF - 1
                               ? __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, fzl), k))
                               : fz64;
          const double R = a.osc >= 0 ? 0.5 * (E[u] - Eprev) : 0.5 * (Eprev - E[u]);
          const double d = (double)fk * inv_tot - R * inv_den;
          cst = fma(tot, d * d, cst);
          Eprev = E[u];
        }
      }
    }
    part += cst;
  }
  wpart[wv][lane] = part;
  __syncthreads();
  if (wv == 0) {
    double t = 0.0;
    for (int q = 0; q < 4; ++q) t += wpart[q][lane];
    w.cost[g * 64 + lane] = t;
  }
}

// ---- 4. Gram matrix of the rocking-curve basis -------------------------------------------------------------------
//
// One wavefront per strong row (rows of a fixed range per wavefront, in order).  Frame k contributes to the Gram
// entries only where some basis function B_kj is non-zero; B_kj = 0 exactly when both frame boundaries saturate every
// component's erf to the same sign, so the accumulation runs over the frames between the first and last live one.
constexpr int GW = 1;                      // wavefronts per workgroup (11.6 KB of LDS each)
__global__ void __launch_bounds__(64 * GW) k_ad_gram(In a, Ws w) {
  __shared__ double A[GW][NF + 1][K + 1];  // boundary erfs E[b][j], then in place A[k][j] = B_kj sqrt(tot), A[k][K] = y
  double (*red)[NE] = (double (*)[NE])&A[0][0][0];   // after the rows: the wavefronts' partial sums
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double (*Aw)[K + 1] = A[wv];
  double *Af = &A[wv][0][0];                // the same storage, flat
  const int gb = blockIdx.x;
  const int n = (int)w.scal[0];
  int ep[4], eq[4];                        // this lane's entries e = lane + 64 i: (p, q), p <= q; q = K: right-hand side
  for (int i = 0; i < 4; ++i) {
    const int e = lane + 64 * i;
    int p = 0, q = 0;
    if (e < K * (K + 1) / 2) {
      int rem = e;
      while (rem >= K - p) {
        rem -= K - p;
        ++p;
      }
      q = p + rem;
    } else if (e < NE) {
      p = e - K * (K + 1) / 2;
      q = K;
    }
    ep[i] = p;
    eq[i] = q;
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (n >= 100) {
    const double sm = w.scal[1];
    int r0, r1;
    row_range(a.n, gb * GW + wv, NG * GW, r0, r1);
    for (int r = r0; r < r1; ++r) {
      if (!w.st[r]) continue;              // uniform over the wavefront
      const double phi = a.phi[r] * (PI / 180.0);
      const double s = sm / fmax(fabs(a.zeta_k[r]), 1e-3);
      const double sw = s * 0.25 * sqrt(2.0);
      const double tot = fmax(w.tot[r], 1e-9), st = sqrt(tot);
      const int iz0 = a.iz0[r];
      // boundary b saturates every component to +1 when (bd - phi) - 3.5 s > ERF_SAT sw, to -1 when
      // (bd - phi) + 3.5 s < -ERF_SAT sw (|RC_C| <= 3.5); frame k is dead (B_kj = 0 for all j) when both its
      // boundaries saturate to the same sign
      const double bdl = bound_deg(a, iz0, lane) * (PI / 180.0) - phi;
      const uint64_t pos_lo = __ballot(bdl - 3.5 * s > ERF_SAT * sw);
      const uint64_t neg_lo = __ballot(bdl + 3.5 * s < -ERF_SAT * sw);
      const double b64 = bound_deg(a, iz0, 64) * (PI / 180.0) - phi, b65 = bound_deg(a, iz0, 65) * (PI / 180.0) - phi;
      const bool pos64 = b64 - 3.5 * s > ERF_SAT * sw, neg64 = b64 + 3.5 * s < -ERF_SAT * sw;
      const bool pos65 = b65 - 3.5 * s > ERF_SAT * sw, neg65 = b65 + 3.5 * s < -ERF_SAT * sw;
      uint64_t dead = (pos_lo & (pos_lo >> 1)) | (neg_lo & (neg_lo >> 1));      // frames 0 .. 62
      if ((((pos_lo >> 63) & 1ull) && pos64) || (((neg_lo >> 63) & 1ull) && neg64)) dead |= 1ull << 63;
      const bool live64 = !((pos64 && pos65) || (neg64 && neg65));
      const uint64_t live = ~dead;
      const int kl = live ? __ffsll((long long)live) - 1 : (live64 ? 64 : NF);
      const int kh = live64 ? 64 : (live ? 63 - __clzll((long long)live) : -1);
      if (kh < kl) continue;
      const int nb = kh - kl + 2;          // boundaries kl .. kh + 1
      for (int u0 = lane; u0 < nb * K; u0 += 4 * 64) {        // four erfs per lane at a time: the reads overlap
        double e[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = min(u0 + 64 * i, nb * K - 1);
          const int b = u / K, j = u - b * K;
          const double bd = bound_deg(a, iz0, kl + b) * (PI / 180.0);
          e[i] = erf_sat(((bd - phi) - RC_C[j] * s) / sw);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = u0 + 64 * i;
          if (u < nb * K) Aw[u / K][u - (u / K) * K] = e[i];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double v[8];                         // (nb - 1) (K + 1) <= 66 x 22 / 64 -> at most 23 per lane: in chunks of 8
      for (int u0 = 0; u0 < (nb - 1) * (K + 1); u0 += 8 * 64) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int u = u0 + lane + 64 * i;
          v[i] = 0.0;
          if (u < (nb - 1) * (K + 1)) {
            const int k = u / (K + 1), j = u - k * (K + 1);
            if (j < K) {
              const double Bk = a.osc >= 0 ? 0.5 * (Aw[k + 1][j] - Aw[k][j]) : 0.5 * (Aw[k][j] - Aw[k + 1][j]);
              v[i] = Bk * st;
            } else {
              v[i] = (double)a.fz[(int64_t)r * NF + kl + k] / tot * st;
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int u = u0 + lane + 64 * i;
          if (u < (nb - 1) * (K + 1)) Af[u] = v[i];      // row-major (K + 1): A[k][j] at k (K + 1) + j
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      for (int k = 0; k < nb - 1; ++k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] += Aw[k][ep[i]] * Aw[k][eq[i]];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __syncthreads();
  for (int i = 0; i < 4; ++i)
    if (lane + 64 * i < NE) red[wv][lane + 64 * i] = acc[i];
  __syncthreads();
  for (int e = threadIdx.x; e < NE; e += 64 * GW) {
    double t = red[0][e];
    for (int q = 1; q < GW; ++q) t += red[q][e];
    w.gram[(int64_t)gb * NE + e] = t;
  }
}

// ---- 5. non-negative least squares and the widths records --------------------------------------------------------

// Gram totals: entry e of each of four quarters of the NG partials summed in order (loads 16 at a time), then the
// quarters in order
__global__ void __launch_bounds__(256) k_ad_gsum(In a, Ws w) {
  __shared__ double red[256];
  if ((int)w.scal[0] < 100) return;
  if (blockIdx.x == NE) {                  // the median beside the Gram totals
    median_select(a, w);
    return;
  }
  const int e = blockIdx.x, t = threadIdx.x;
  // entry e: thread t sums the partials t, t + 256, ... in order, then a fixed tree over the threads
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NG / 256; ++i) s += w.gram[(int64_t)(t + 256 * i) * NE + e];
  s = block_sum256(s, red);
  if (t == 0) w.gtot[e] = s;
}

// Lawson-Hanson active-set NNLS in Gram form (min w^T G w - 2 w^T h, w >= 0) on one wavefront.
//
// Lane j holds row j of G and x_j.  The active block G[P,P] is kept as a Cholesky factor in insertion order: lane i
// holds row i (a) and column i (ct) of L, and z = L^-1 h[P].  Adding a variable appends one row (a forward
// substitution); removing variables refactors by re-appending the remaining ones in order.  Every index into a
// per-lane array is static; broadcasts are readlanes.
// compile-time loop: f(IC<0>{}), ..., f(IC<N - 1>{}) in order (indices are constants, so per-lane arrays stay in
// registers)
template <int I>
struct IC {
  static constexpr int value = I;
};
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (N > 0) {
    static_for<N - 1>(f);
    f(IC<N - 1>{});
  }
}
template <int N, typename F>
__device__ __forceinline__ void static_for_down(F &&f) {     // f(IC<N - 1>{}), ..., f(IC<0>{})
  if constexpr (N > 0) {
    f(IC<N - 1>{});
    static_for_down<N - 1>(f);
  }
}

struct Nnls {
  double a[K], ct[K];
  double invd;     // lane i < m: 1 / L_ii
  double zv;       // lane i < m: z_i
  int ordv;        // lane i < m: variable at position i
  int posv;        // lane j: position of variable j, or -1
  int m;
};

__device__ __forceinline__ void nnls_reset(Nnls &S) {
  static_for<K>([&](auto k) { S.a[k.value] = S.ct[k.value] = 0.0; });
  S.invd = 0.0;
  S.zv = 0.0;
  S.ordv = -1;
  S.posv = -1;
  S.m = 0;
}

// append variable t (uniform); false (state unchanged) when G[P+t, P+t] is not numerically positive definite.
// The substitutions multiply by the stored reciprocal diagonal and use fused multiply-adds: each step of a triangular
// solve is a readlane, a multiply and one fma.
__device__ __forceinline__ bool nnls_append(Nnls &S, const double *Gs, const double *hs, int t) {
  const int lane = threadIdx.x & 63;
  const int m = S.m;
  double b = lane < m ? Gs[S.ordv * K + t] : 0.0;
  // L l = G[P, t], column-oriented: after step k lane k holds l_k
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < m) {
      const double lk = rld(b, k) * rld(S.invd, k);
      if (lane == k) b = lk;
      else if (lane > k) b = fma(-S.a[k], lk, b);
    }
  });
  double d2 = Gs[t * K + t], zm = hs[t];
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < m) {
      const double lk = rld(b, k);
      d2 = fma(-lk, lk, d2);
      zm = fma(-lk, rld(S.zv, k), zm);
    }
  });
  if (!(d2 > 0)) return false;
  const double d = sqrt(d2);
  const double id = 1.0 / d;
  zm = zm * id;
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    // unconditional stores of selected values keep the per-lane arrays in registers
    if (k < m) {
      const double lk = rld(b, k);
      S.a[k] = lane == m ? lk : S.a[k];
    } else if (k == m) {
      S.a[k] = lane == m ? d : S.a[k];
      S.ct[k] = lane < m ? b : (lane == m ? d : S.ct[k]);
    }
  });
  if (lane == m) {
    S.zv = zm;
    S.ordv = t;
    S.invd = id;
  }
  if (lane == t) S.posv = m;
  S.m = m + 1;
  return true;
}

// Factor an ordered set at once (lane i: ordl = the variable at position i, -1 past the m positions): right-looking
// Cholesky with row i of the working matrix in lane i, so the dependent chain is m pivots instead of the m^2 / 2 steps
// of m appends.  Fills S as the appends would (up to rounding).  Returns false, S reset, when a pivot is not
// numerically positive; the caller then appends the variables one at a time, which skips such a variable.  Each
// entry is computed by the same operations in the same order as by the appends (the same fma chains over the same
// products), so the factor, z and the solution are bit-identical to appending the set in order.
// X (K x K doubles of LDS) is scratch for the transposition.
__device__ __forceinline__ bool nnls_factor(Nnls &S, const double *Gs, const double *hs, int ordl, int m, double *X,
                                          int *pos) {
  const int lane = threadIdx.x & 63;
  const bool in = lane < m;
  const int oi = in ? ordl : 0;
  double a[K];
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const int ok = __builtin_amdgcn_readlane(ordl, k < m ? k : 0);
    a[k] = (in && k < m && ok >= 0) ? Gs[oi * K + (ok >= 0 ? ok : 0)] : 0.0;
  });
  double invd = 0.0;
  bool good = true;
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < m && good) {
      const double piv = rld(a[k], k);
      if (!(piv > 0)) {
        good = false;
      } else {
        const double d = sqrt(piv), id = 1.0 / d;
        if (lane == k) {
          a[k] = d;
          invd = id;
        } else if (lane > k) {
          a[k] = a[k] * id;
        }
        static_for<K>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if constexpr (j > k) {
            if (j < m) {
              const double ljk = rld(a[k], j);
              if (lane >= j) a[j] = fma(-a[k], ljk, a[j]);
            }
          }
        });
      }
    }
  });
  nnls_reset(S);
  if (!good) return false;
  // rows of L (zero above the diagonal and past m), their transpose into ct through LDS
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    S.a[k] = (in && k <= lane) ? a[k] : 0.0;
  });
  if (lane < K) {
    static_for<K>([&](auto kc) { X[lane * K + kc.value] = S.a[kc.value]; });
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    S.ct[k] = (lane < K && lane <= k && k < m) ? X[k * K + lane] : 0.0;
  });
  // z = L^-1 h[P], column-oriented in position order (the appends' order of operations)
  double z = in ? hs[oi] : 0.0;
  static_for<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < m) {
      const double zk = rld(z, k) * rld(invd, k);
      if (lane == k) z = zk;
      else if (lane > k) z = fma(-S.a[k], zk, z);
    }
  });
  // positions of the variables
  if (lane < K) pos[lane] = -1;
  __syncthreads();
  if (in) pos[oi] = lane;
  __syncthreads();
  S.zv = in ? z : 0.0;
  S.invd = in ? invd : 0.0;
  S.ordv = in ? ordl : -1;
  S.posv = lane < K ? pos[lane] : -1;
  S.m = m;
  __syncthreads();
  return true;
}

// L^T s = z; returns s_j on lane j (0 for inactive variables)
__device__ __forceinline__ double nnls_solve(const Nnls &S) {
  const int lane = threadIdx.x & 63;
  const int m = S.m;
  double c = lane < m ? S.zv : 0.0;
  static_for_down<K>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k < m) {
      const double sk = rld(c, k) * rld(S.invd, k);
      if (lane == k) c = sk;
      else if (lane < k) c = fma(-S.ct[k], sk, c);
    }
  });
  const double sv = __shfl(c, S.posv < 0 ? 0 : S.posv, 64);
  return S.posv < 0 ? 0.0 : sv;
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) k_ad_solve(In a, Ws w, const xpg_widths *in, xpg_widths *out2,
                                                 xpg_widths *outn) {
  __shared__ double Gs[K * K], hs[K];
  const int n = (int)w.scal[0];
  const int lane = threadIdx.x;
  if (n < 100) {
    if (lane == 0) {
      xpg_widths v = *in;
      v.n_strong = n;
      v.sd_fit = 0.0;
      v.sm_fit = 0.0;
      *out2 = v;
      *outn = v;
    }
    return;
  }
  for (int u = lane; u < K * K; u += 64) {
    const int i = u / K, k = u - i * K;
    const int p = i < k ? i : k, q = i < k ? k : i;
    Gs[u] = w.gtot[p * K - p * (p - 1) / 2 + (q - p)];
  }
  if (lane < K) hs[lane] = w.gtot[K * (K + 1) / 2 + lane];
  __syncthreads();
  double grow[K];
  const int il = lane < K ? lane : 0;
  static_for<K>([&](auto k) { grow[k.value] = lane < K ? Gs[il * K + k.value] : 0.0; });
  const double hl = lane < K ? hs[il] : 0.0;
  double gmax = 0.0;
  static_for<K>([&](auto k) { gmax = fmax(gmax, fabs(grow[k.value])); });
  for (int o = 32; o > 0; o >>= 1) gmax = fmax(gmax, __shfl_xor(gmax, o, 64));
  const double tol = 10.0 * 2.220446049250313e-16 * gmax * K;
  __shared__ double Xs[K * K];
  __shared__ int ords[64], poss[K];
  Nnls S;
  nnls_reset(S);
  double x = 0.0;                          // lane j: x_j
  int n_outer = 0, n_inner = 0;
  // lane i: the i-th set bit of `bits` (variables in index order), -1 past them
  auto ordered_bits = [&](uint32_t bits) {
    ords[lane] = -1;
    __syncthreads();
    if (lane < K && ((bits >> lane) & 1u)) ords[__popc(bits & ((1u << lane) - 1u))] = lane;
    __syncthreads();
    const int o = ords[lane];
    __syncthreads();
    return o;
  };
  // lane i: the i-th kept variable of the order held in `ordv` (positions below mo, keep: bit per variable)
  auto kept_order = [&](int ordv, int mo, uint64_t keepv) {
    const bool kp = lane < mo && ordv >= 0 && ((keepv >> ordv) & 1ull);
    const uint64_t kb = __ballot(kp);
    ords[lane] = -1;
    __syncthreads();
    if (kp) ords[__popcll(kb & ((1ull << lane) - 1ull))] = ordv;
    __syncthreads();
    const int o = ords[lane];
    __syncthreads();
    return o;
  };
  // factor the ordered set at once; when a pivot is not positive, append one at a time as before (a variable that
  // cannot be appended is left out, and with zero_failed its x is zeroed)
  auto refactor = [&](int ordl, int mnew, bool zero_failed) {
    if (nnls_factor(S, Gs, hs, ordl, mnew, Xs, poss)) return;
    nnls_reset(S);
    for (int i = 0; i < mnew; ++i) {
      const int v = __builtin_amdgcn_readlane(ordl, i);
      if (!nnls_append(S, Gs, hs, v) && zero_failed && lane == v) x = 0.0;
    }
  };
  // Warm start from the previous fit's positive set: append it, drop the variables whose unconstrained solution is not
  // positive until all are, and continue from that point (x_P the minimiser on P, x >= 0: the invariant of the
  // outer loop).  The minimiser is unique for a positive definite G, so only the iteration count depends on it.
  {
    const uint32_t warm = in->active & ((1u << K) - 1u);
    refactor(ordered_bits(warm), __popc(warm), false);
    for (int pass = 0; pass < K && S.m > 0; ++pass) {
      ++n_inner;
      const double sj = nnls_solve(S);
      const bool act = S.posv >= 0;
      const bool bad = act && !(sj > tol);
      if (!__ballot(bad)) {
        x = act ? sj : 0.0;
        break;
      }
      const uint64_t keep = __ballot(lane < K && act && !bad);
      const int mo = S.m;
      refactor(kept_order(S.ordv, mo, keep), __popcll(keep), false);
    }
  }
  for (int outer = 0; outer < 3 * K; ++outer) {
    ++n_outer;
    double g = hl;
    static_for<K>([&](auto k) { g = fma(-grow[k.value], rld(x, k.value), g); });
    // the first variable outside P with the largest gradient above tol
    const bool cand = lane < K && S.posv < 0 && g > tol;
    double gb = cand ? g : -INFINITY;
    for (int o = 32; o > 0; o >>= 1) gb = fmax(gb, __shfl_xor(gb, o, 64));
    const uint64_t hit = __ballot(cand && g == gb);
    if (!hit) break;
    const int t = __ffsll((long long)hit) - 1;
    if (!nnls_append(S, Gs, hs, t)) break;   // the same t would be chosen again with x unchanged
    for (int inner = 0; inner < 3 * K; ++inner) {
      ++n_inner;
      const double sj = nnls_solve(S);
      const bool act = S.posv >= 0;
      if (!__ballot(act && sj <= 0)) {
        x = act ? sj : 0.0;
        break;
      }
      double al = (act && sj <= 0) ? x / (x - sj) : INFINITY;
      for (int o = 32; o > 0; o >>= 1) al = fmin(al, __shfl_xor(al, o, 64));
      x = x + al * (sj - x);
      const bool drop = act && x <= tol;
      if (drop) x = 0.0;
      if (!__ballot(drop)) continue;
      // refactor: the remaining variables in their previous order
      const int mo = S.m;
      const bool keep_l = lane < K && S.posv >= 0 && !drop;
      const uint64_t keep = __ballot(keep_l);
      refactor(kept_order(S.ordv, mo, keep), __popcll(keep), true);
    }
  }
  double xs = x;
  double sw = 0.0;
#pragma unroll
  for (int j = 0; j < K; ++j) sw += rld(xs, j);
  sw = fmax(sw, 1e-9);
  // default weights: exp(-c^2 / (2 (1 - RC_W^2))) normalised, as float32 values
  double s0 = 0.0;
#pragma unroll
  for (int j = 0; j < K; ++j) s0 += exp(-0.5 * RC_C[j] * RC_C[j] / (1 - 0.25 * 0.25));
  const double lam = 200.0 / (double)n;
  float wj = 0.f;
  if (lane < K) {
    const double d = (double)(float)(exp(-0.5 * RC_C[lane] * RC_C[lane] / (1 - 0.25 * 0.25)) / s0);
    wj = (float)((xs / sw + lam * d) / (1 + lam));
  }
  if (lane < K) {
    out2->rcw[lane] = wj;
    outn->rcw[lane] = wj;
  }
  const uint32_t act_out = (uint32_t)__ballot(lane < K && xs > 0.0);
  if (lane == 0) {
    w.scal[5] = (double)n_outer;           // iteration counts of the last solve (diagnostics)
    w.scal[6] = (double)n_inner;
    out2->active = act_out;
    outn->active = act_out;
    out2->sd = in->sd;
    out2->sm = w.scal[1];
    out2->n_strong = n;
    out2->sm_fit = w.scal[2];
    out2->sd_fit = w.scal[4];
    outn->sd = w.scal[3];
    outn->sm = w.scal[1];
    outn->n_strong = n;
    outn->sm_fit = w.scal[2];
    outn->sd_fit = w.scal[4];
  }
}

}  // namespace

extern "C" {

int64_t xpg_adapt_workspace_bytes(int32_t n) {
  Ws w = carve(nullptr, n);
  return (int64_t)((char *)w.kc - (char *)nullptr) + 8 * (int64_t)(n > 0 ? n : 1) + 256;
}

int xpg_adapt_block(int32_t n, const double *isum, const double *vsum, const int32_t *n_over, const double *frac_rec,
                    const double *m_d, const float *fz, const double *phi_deg, const double *zeta_k,
                    const int32_t *iz0, const int32_t *fb, const double *Xlen, const double *cosa, double sd_run,
                    double sm_run, double phi0_deg, double osc_deg, double qx, const xpg_widths *in,
                    xpg_widths *out_pass2, xpg_widths *out_next, void *workspace, int64_t workspace_bytes,
                    void *stream) {
  if (n < 0 || !in || !out_pass2 || !out_next || !workspace) return XPG_ERR_ARG;
  if (n > 0 && (!isum || !vsum || !n_over || !frac_rec || !m_d || !fz || !phi_deg || !zeta_k || !iz0 || !fb ||
                !Xlen || !cosa))
    return XPG_ERR_ARG;
  if (workspace_bytes < xpg_adapt_workspace_bytes(n)) return XPG_ERR_ARG;
  In a;
  a.n = n;
  a.isum = isum;
  a.vsum = vsum;
  a.frac_rec = frac_rec;
  a.m_d = m_d;
  a.phi = phi_deg;
  a.zeta_k = zeta_k;
  a.Xlen = Xlen;
  a.cosa = cosa;
  a.n_over = n_over;
  a.iz0 = iz0;
  a.fb = fb;
  a.fz = fz;
  a.sd_run = sd_run;
  a.sm_run = sm_run;
  a.phi0 = phi0_deg;
  a.osc = osc_deg;
  a.qx = qx;
  Ws w = carve(workspace, n);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ad_prep, dim3(w.nprep), dim3(256), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_grid, dim3(NGW), dim3(256), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_finish, dim3(1), dim3(256), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_gram, dim3(NG), dim3(64 * GW), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_gsum, dim3(NE + 1), dim3(256), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_solve, dim3(1), dim3(64), 0, st, a, w, in, out_pass2, out_next);
  return hipGetLastError() == hipSuccess ? XPG_OK : XPG_ERR_HIP;
}

}  // extern "C"
