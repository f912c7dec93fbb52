// Adaptive widths per phi block on the device (xpg_adapt_block, include/xpg_integrate.h): the small fits of
// xpg/integrate.py:404-461 as a chain of five launches on the caller's stream, so the per-block chain
// pass 1 -> fits -> pass 2 -> next block runs without a host round trip.
//
//   k_ad_select  (1 x 1024)  strong reflections in row order (block-wide scan), their fz totals and sigma_D terms
//   k_ad_grid    (61 x 256)  cost of each sigma_M of the 61-point log grid (_fit_rocking_width, :1130-1154)
//   k_ad_finish  (1 x 1024)  argmin -> sigma_M fit and shrinkage; median of the sigma_D terms (bitonic sort)
//   k_ad_gram    (NG x 256)  Gram matrix A^T A and A^T y of the rocking-curve basis (rc_basis_np, :437-443)
//   k_ad_solve   (1 x 256)   fixed-order reduction, non-negative least squares on the Gram form (_nnls_gram,
//                            :1095-1104), regularisation towards the Gaussian weights, the widths records
//
// All sums are float64 in fixed orders: results do not depend on scheduling.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "xpg_integrate.h"

namespace {

constexpr int K = XPG_RC_K;               // 21 mixture components
constexpr int NF = 65;                     // frames of a per-frame profile row (max(DZ_BUCKETS))
constexpr int NG = 64;                     // workgroups of the Gram accumulation
constexpr int NE = K * (K + 1) / 2 + K;    // unique Gram entries + right-hand side
constexpr int NGRID = 61;
constexpr double PI = 3.141592653589793;

// rc_default_weights (float32 values of exp(-c^2 / (2 (1 - 0.25^2))) normalised)
__device__ __constant__ static const double RC_C[K] = {
    -3.5, -3.15, -2.8, -2.4499999999999997, -2.1, -1.75, -1.4, -1.0499999999999998, -0.7, -0.34999999999999964, 0.0,
    0.35000000000000053, 0.7000000000000002, 1.0500000000000007, 1.4000000000000004, 1.75, 2.1000000000000005,
    2.4500000000000006, 2.8000000000000007, 3.1500000000000004, 3.5};

struct Ws {
  int32_t *idx;
  double *tot, *md, *sortbuf, *cost, *gram, *scal;   // scal: n, sm_new, sm_fit, sd_new, sd_fit
  int32_t npow;
};

__host__ __device__ inline int32_t pow2ceil(int32_t n) {
  int32_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

__host__ __device__ inline Ws carve(void *base, int32_t n) {
  Ws w;
  char *p = (char *)base;
  auto take = [&](size_t bytes) {
    char *q = p;
    p += (bytes + 255) & ~(size_t)255;
    return q;
  };
  w.npow = pow2ceil(n > 0 ? n : 1);
  w.idx = (int32_t *)take(4 * (size_t)(n > 0 ? n : 1));
  w.tot = (double *)take(8 * (size_t)(n > 0 ? n : 1));
  w.md = (double *)take(8 * (size_t)(n > 0 ? n : 1));
  w.sortbuf = (double *)take(8 * (size_t)w.npow);
  w.cost = (double *)take(8 * NGRID);
  w.gram = (double *)take(8 * (size_t)NG * NE);
  w.scal = (double *)take(8 * 8);
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

// ---- 1. strong reflections in row order ----------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_ad_select(In a, Ws w) {
  __shared__ int32_t cnt[1024];
  const int t = threadIdx.x, T = blockDim.x;
  const int chunk = (a.n + T - 1) / T;
  const int r0 = min(t * chunk, a.n), r1 = min(r0 + chunk, a.n);
  int c = 0;
  for (int r = r0; r < r1; ++r) {
    const bool st = (a.isum[r] > 0) && (a.isum[r] / sqrt(fmax(a.vsum[r], 1e-6)) > 15.0) && (a.n_over[r] == 0) &&
                    (a.frac_rec[r] > 0.98) && (a.fb[r] == 1);
    c += st;
  }
  cnt[t] = c;
  __syncthreads();
  // inclusive scan (Hillis-Steele)
  for (int off = 1; off < T; off <<= 1) {
    int v = t >= off ? cnt[t - off] : 0;
    __syncthreads();
    cnt[t] += v;
    __syncthreads();
  }
  int o = cnt[t] - c;
  const double t2 = 1.0 - (9.0 / 2) * exp(-9.0 / 2) / (1 - exp(-9.0 / 2));
  for (int r = r0; r < r1; ++r) {
    const bool st = (a.isum[r] > 0) && (a.isum[r] / sqrt(fmax(a.vsum[r], 1e-6)) > 15.0) && (a.n_over[r] == 0) &&
                    (a.frac_rec[r] > 0.98) && (a.fb[r] == 1);
    if (!st) continue;
    w.idx[o] = r;
    double s = 0.0;
    for (int k = 0; k < NF; ++k) s += (double)a.fz[(int64_t)r * NF + k];
    w.tot[o] = s;
    const double q = a.qx / a.Xlen[r];
    w.md[o] = a.m_d[r] / t2 - (q * q) * (1 + a.cosa[r] * a.cosa[r]) / 24.0;
    ++o;
  }
  if (t == T - 1) w.scal[0] = (double)cnt[T - 1];
}

// ---- 2. grid of sigma_M candidates -------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ad_grid(In a, Ws w) {
  __shared__ double red[256];
  const int n = (int)w.scal[0];
  const int c = blockIdx.x;
  double part = 0.0;
  if (n >= 100) {
    const double guess = a.sm_run * (180.0 / PI);
    const double lo_l = log(0.5), hi_l = log(2.0), step = (hi_l - lo_l) / (NGRID - 1);
    const double smv = guess * exp(c == NGRID - 1 ? hi_l : (double)c * step + lo_l);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double tot = w.tot[i];
      if (!(tot > 0)) continue;
      const int r = w.idx[i];
      const double phi = a.phi[r], zeta = fabs(a.zeta_k[r]);
      const double s = smv / zeta * sqrt(2.0);
      const float *fz = a.fz + (int64_t)r * NF;
      double R[NF];
      double Rs = 0.0;
      double Eprev = erf((bound_deg(a, a.iz0[r], 0) - phi) / s);
      for (int k = 0; k < NF; ++k) {
        const double E = erf((bound_deg(a, a.iz0[r], k + 1) - phi) / s);
        const double v = a.osc >= 0 ? 0.5 * (E - Eprev) : 0.5 * (Eprev - E);
        R[k] = v;
        Rs = Rs + v;
        Eprev = E;
      }
      const double den = fmax(Rs, 1e-9);
      double cst = 0.0;
      for (int k = 0; k < NF; ++k) {
        const double d = (double)fz[k] / tot - R[k] / den;
        cst += tot * (d * d);
      }
      part += cst;
    }
  }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) w.cost[c] = red[0];
}

// ---- 3. sigma_M from the grid, sigma_D from the median ----------------------------------------------------------
__global__ void __launch_bounds__(1024) k_ad_finish(In a, Ws w) {
  const int n = (int)w.scal[0];
  if (n < 100) return;
  const int t = threadIdx.x, T = blockDim.x;
  if (t == 0) {
    const double guess = a.sm_run * (180.0 / PI);
    const double lo_l = log(0.5), hi_l = log(2.0), step = (hi_l - lo_l) / (NGRID - 1);
    double best = INFINITY, bsm = guess;
    for (int c = 0; c < NGRID; ++c) {
      if (w.cost[c] < best) {
        best = w.cost[c];
        bsm = guess * exp(c == NGRID - 1 ? hi_l : (double)c * step + lo_l);
      }
    }
    const double sm_fit = bsm * (PI / 180.0);
    const double sm_new = exp(((double)n * log(sm_fit) + 300.0 * log(a.sm_run)) / ((double)n + 300.0));
    w.scal[1] = sm_new;
    w.scal[2] = sm_fit;
  }
  // bitonic sort of the sigma_D terms (padded with +inf)
  const int N = w.npow;
  for (int i = t; i < N; i += T) w.sortbuf[i] = i < n ? w.md[i] : INFINITY;
  __syncthreads();
  for (int k = 2; k <= N; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < N; i += T) {
        const int l = i ^ j;
        if (l > i) {
          const double x = w.sortbuf[i], y = w.sortbuf[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            w.sortbuf[i] = y;
            w.sortbuf[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  if (t == 0) {
    const double med = (n & 1) ? w.sortbuf[n / 2] : (w.sortbuf[n / 2 - 1] + w.sortbuf[n / 2]) / 2.0;
    double sd_fit = sqrt(fmax(med, 1e-12));
    sd_fit = fmin(fmax(sd_fit, 0.5 * a.sd_run), 2.0 * a.sd_run);
    w.scal[3] = exp(((double)n * log(sd_fit) + 300.0 * log(a.sd_run)) / ((double)n + 300.0));
    w.scal[4] = sd_fit;
  }
}

// ---- 4. Gram matrix of the rocking-curve basis -------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ad_gram(In a, Ws w) {
  __shared__ double A[NF][K + 1];          // A[k][j] = B_kj sqrt(tot), A[k][K] = y_k
  __shared__ double E[NF + 1][K];
  const int n = (int)w.scal[0];
  double acc = 0.0;
  const int e = threadIdx.x;               // entry: (p, q) with p <= q, then the right-hand side
  int p = 0, q = 0;
  if (e < NE) {
    if (e < K * (K + 1) / 2) {
      int rem = e;
      p = 0;
      while (rem >= K - p) {
        rem -= K - p;
        ++p;
      }
      q = p + rem;
    } else {
      p = e - K * (K + 1) / 2;
      q = K;
    }
  }
  if (n >= 100) {
    const double sm = w.scal[1];
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
      const int r = w.idx[i];
      const double phi = a.phi[r] * (PI / 180.0);
      const double s = sm / fmax(fabs(a.zeta_k[r]), 1e-3);
      const double sw = s * 0.25 * sqrt(2.0);
      const double tot = fmax(w.tot[i], 1e-9), st = sqrt(tot);
      // erf at the 66 frame boundaries for each component (lo / hi of frame k share boundaries)
      for (int u = threadIdx.x; u < (NF + 1) * K; u += blockDim.x) {
        const int b = u / K, j = u - b * K;
        const double bd = bound_deg(a, a.iz0[r], b) * (PI / 180.0);
        E[b][j] = erf(((bd - phi) - RC_C[j] * s) / sw);
      }
      __syncthreads();
      for (int u = threadIdx.x; u < NF * (K + 1); u += blockDim.x) {
        const int k = u / (K + 1), j = u - k * (K + 1);
        if (j < K) {
          const double Bk = a.osc >= 0 ? 0.5 * (E[k + 1][j] - E[k][j]) : 0.5 * (E[k][j] - E[k + 1][j]);
          A[k][j] = Bk * st;
        } else {
          A[k][K] = (double)a.fz[(int64_t)r * NF + k] / tot * st;
        }
      }
      __syncthreads();
      if (e < NE)
        for (int k = 0; k < NF; ++k) acc += A[k][p] * A[k][q];
      __syncthreads();
    }
  }
  if (e < NE) w.gram[(int64_t)blockIdx.x * NE + e] = acc;
}

// ---- 5. non-negative least squares and the widths records --------------------------------------------------------

// solve G[P,P] s = h[P] by Cholesky (|P| <= 21); returns false when not positive definite
__device__ bool chol_solve(const double G[K][K], const double *h, const bool *inP, double *s) {
  int idx[K], m = 0;
  for (int j = 0; j < K; ++j)
    if (inP[j]) idx[m++] = j;
  double L[K][K];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j) {
      double v = G[idx[i]][idx[j]];
      for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(v > 0)) return false;
        L[i][i] = sqrt(v);
      } else {
        L[i][j] = v / L[j][j];
      }
    }
  double z[K];
  for (int i = 0; i < m; ++i) {
    double v = h[idx[i]];
    for (int k = 0; k < i; ++k) v -= L[i][k] * z[k];
    z[i] = v / L[i][i];
  }
  for (int j = 0; j < K; ++j) s[j] = 0.0;
  for (int i = m - 1; i >= 0; --i) {
    double v = z[i];
    for (int k = i + 1; k < m; ++k) v -= L[k][i] * s[idx[k]];
    s[idx[i]] = v / L[i][i];
  }
  return true;
}

__global__ void __launch_bounds__(256) k_ad_solve(In a, Ws w, const xpg_widths *in, xpg_widths *out2,
                                                  xpg_widths *outn) {
  __shared__ double tot[NE];
  const int n = (int)w.scal[0];
  if (n < 100) {
    if (threadIdx.x == 0) {
      xpg_widths v = *in;
      v.n_strong = n;
      v.sd_fit = 0.0;
      v.sm_fit = 0.0;
      *out2 = v;
      *outn = v;
    }
    return;
  }
  for (int e = threadIdx.x; e < NE; e += blockDim.x) {
    double s = 0.0;
    for (int g = 0; g < NG; ++g) s += w.gram[(int64_t)g * NE + e];
    tot[e] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double G[K][K], h[K];
  {
    int e = 0;
    for (int p = 0; p < K; ++p)
      for (int q = p; q < K; ++q, ++e) G[p][q] = G[q][p] = tot[e];
    for (int p = 0; p < K; ++p) h[p] = tot[K * (K + 1) / 2 + p];
  }
  // active-set NNLS (Lawson-Hanson, Gram form): min w^T G w - 2 w^T h, w >= 0
  double gmax = 0.0;
  for (int p = 0; p < K; ++p)
    for (int q = 0; q < K; ++q) gmax = fmax(gmax, fabs(G[p][q]));
  const double tol = 10.0 * 2.220446049250313e-16 * gmax * K;
  bool inP[K];
  double x[K], s[K], g[K];
  for (int j = 0; j < K; ++j) {
    inP[j] = false;
    x[j] = 0.0;
  }
  for (int outer = 0; outer < 3 * K; ++outer) {
    for (int j = 0; j < K; ++j) {
      double v = h[j];
      for (int k = 0; k < K; ++k) v -= G[j][k] * x[k];
      g[j] = v;
    }
    int t = -1;
    double best = tol;
    for (int j = 0; j < K; ++j)
      if (!inP[j] && g[j] > best) {
        best = g[j];
        t = j;
      }
    if (t < 0) break;
    inP[t] = true;
    for (int inner = 0; inner < 3 * K; ++inner) {
      if (!chol_solve(G, h, inP, s)) {
        inP[t] = false;
        break;
      }
      bool allpos = true;
      for (int j = 0; j < K; ++j)
        if (inP[j] && s[j] <= 0) allpos = false;
      if (allpos) {
        for (int j = 0; j < K; ++j) x[j] = s[j];
        break;
      }
      double alpha = INFINITY;
      for (int j = 0; j < K; ++j)
        if (inP[j] && s[j] <= 0) alpha = fmin(alpha, x[j] / (x[j] - s[j]));
      for (int j = 0; j < K; ++j) {
        x[j] = x[j] + alpha * (s[j] - x[j]);
        if (inP[j] && x[j] <= tol) {
          inP[j] = false;
          x[j] = 0.0;
        }
      }
    }
  }
  double sw = 0.0;
  for (int j = 0; j < K; ++j) sw += x[j];
  sw = fmax(sw, 1e-9);
  // default weights: exp(-c^2 / (2 (1 - RC_W^2))) normalised, as float32 values
  double w0[K], s0 = 0.0;
  for (int j = 0; j < K; ++j) {
    w0[j] = exp(-0.5 * RC_C[j] * RC_C[j] / (1 - 0.25 * 0.25));
    s0 += w0[j];
  }
  const double lam = 200.0 / (double)n;
  xpg_widths o2 = *in, on;
  for (int j = 0; j < K; ++j) {
    const double d = (double)(float)(w0[j] / s0);
    o2.rcw[j] = (float)((x[j] / sw + lam * d) / (1 + lam));
  }
  o2.sm = w.scal[1];
  o2.n_strong = n;
  o2.sm_fit = w.scal[2];
  o2.sd_fit = w.scal[4];
  on = o2;
  on.sd = w.scal[3];
  *out2 = o2;
  *outn = on;
}

}  // namespace

extern "C" {

int64_t xpg_adapt_workspace_bytes(int32_t n) {
  Ws w = carve(nullptr, n);
  return (int64_t)((char *)w.scal - (char *)nullptr) + 256;
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
  hipLaunchKernelGGL(k_ad_select, dim3(1), dim3(1024), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_grid, dim3(NGRID), dim3(256), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_finish, dim3(1), dim3(1024), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_gram, dim3(NG), dim3(256), 0, st, a, w);
  hipLaunchKernelGGL(k_ad_solve, dim3(1), dim3(256), 0, st, a, w, in, out_pass2, out_next);
  return hipGetLastError() == hipSuccess ? XPG_OK : XPG_ERR_HIP;
}

}  // extern "C"
