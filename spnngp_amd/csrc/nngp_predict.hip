// nngp_predict.hip — kriging kernels.
//
// Response model, src/rNNGPPredict.cpp:139-183: for every new site i and
// posterior sample s, with the m fitted neighbors N of i,
//   C = s2_s R(N,N; phi_s) + t2_s I,  c = s2_s R(N,i; phi_s),  w = C^-1 c,
//   y0[s*q+i] = x0_i.beta_s + sum_k w_k (y_k - x_k.beta_s)
//               + sqrt(s2_s + t2_s - w.c) z[i*S+s].
// The neighbor geometry does not depend on s: one block owns one site, puts
// its neighbors' coordinates, pairwise distances, X rows and y in LDS once,
// and each lane solves the system of one sample.
//
// Latent model, src/sNNGPPredict.cpp:163-205: the same systems without the
// nugget, weights applied to the posterior samples of w at the neighbors,
//   w0[s*q+i] = sum_k w_k wS[s*n + N_k] + sqrt(s2_s - w.c) z[i*S+s],
//   y0[s*q+i] = sqrt(t2_s) z2[i*S+s] + x0_i.beta_s + w0[s*q+i].
//
// Conjugate model, src/cNNGP.cpp:255-299: one system per held-out site
// (one lane each) for the grid point's (phi, alpha):
//   y0 = x0.beta + w.(y_N - X_N beta),  u = x0 - X_N^T w,
//   var = b (u^T Vb u + 1 + alpha - w.c0) / (a - 1).
#include "nngp_kernels.hpp"
#include "nngp_wave.hpp"

namespace nngp {

constexpr int kPredBlock = 256;
constexpr int kPredMaxM = 16;
constexpr int kPredMaxP = kMaxP;

// correlation from a distance (not squared) with unit scale; ip = 1/phi,
// phs = phi times the unit record's tsc (the base-2 scale of s2_exp2)
template <int COV>
__device__ __forceinline__ double corr_d(double D, double phi, double ip,
                                         double phs, const CovArgs &u) {
  if constexpr (COV == 0) {
    return s2_exp2(-phs * D, u);  // exp(-phi D) = 2^(-phi D / ln 2)
  } else if constexpr (COV == 1) {
    const double t = phi * D;
    const double v = 1.0 - 1.5 * t + 0.5 * (t * t * t);
    return (D > 0.0 && D <= ip) ? v : (D >= ip ? 0.0 : 1.0);
  } else {
    const double t = phs * D;  // phi D / sqrt(ln 2)
    return s2_exp2(-(t * t), u);
  }
}

// The new site a block works on, with the arrays of its request.
struct SiteRef {
  const double2 *xy0;
  const double *X0;
  const int *nn0;
  double *y0;
  const double *z;  // nullptr: z from philox_normal(seed, i, s)
  int *err;
  unsigned long long seed;
  int q, i;
};

__device__ __forceinline__ SiteRef pred_site(const PredArgs &A, int b) {
  SiteRef r;
  if (A.segs == nullptr) {
    r.xy0 = A.xy0;
    r.X0 = A.X0;
    r.nn0 = A.nn0;
    r.y0 = A.y0;
    r.z = A.z;
    r.err = A.err;
    r.seed = 0;
    r.q = A.q;
    r.i = A.site0 + b;
    return r;
  }
  int lo = 0, hi = A.nseg - 1;  // last run with block0 <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A.segs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const PredSeg &g = A.segs[lo];
  r.xy0 = g.xy0;
  r.X0 = g.X0;
  r.nn0 = g.nn0;
  r.y0 = g.y0;
  r.z = nullptr;
  r.err = g.err;
  r.seed = g.seed;
  r.q = g.q;
  r.i = g.first + (b - g.block0);
  return r;
}

__device__ __forceinline__ double pred_z(const SiteRef &R, int S, int s) {
  return R.z ? R.z[(size_t)R.i * S + s]
             : philox_normal(R.seed, (uint32_t)R.i, (uint32_t)s);
}

// One block = one site x kPredBlock samples.  Grid: (nsite, ceil(S/256)).
// M >= m; rows m..M-1 are padding at infinite distance (exactly zero
// correlation), which leaves the real block's arithmetic unchanged.
template <int M, int COV>
__global__ __launch_bounds__(kPredBlock) void k_pred_lane(PredArgs A) {
  __shared__ double sD[M * (M - 1) / 2 + 1];  // pairwise distances (l < k)
  __shared__ double sD0[M];                   // site-neighbor distances
  __shared__ double sY[M];
  __shared__ double sX[M * kPredMaxP];        // X rows of the neighbors
  __shared__ double sX0[kPredMaxP];
  __shared__ double sPx[M], sPy[M];
  __shared__ int sId[M];
  const SiteRef R = pred_site(A, (int)blockIdx.x);
  const int i = R.i;
  const int t = (int)threadIdx.x;
  const int m = A.m;
  if (t < M) {
    double px = 0.0, py = 0.0, yy = 0.0;
    int id = 0;
    if (t < m) {
      id = R.nn0[(size_t)t * R.q + i];
      const double2 pt = A.xy[id];
      px = pt.x;
      py = pt.y;
      yy = A.y ? A.y[id] : 0.0;  // no y in the latent form
      for (int j = 0; j < A.p; j++)
        sX[t * kPredMaxP + j] = A.X[(size_t)j * A.n + id];
    } else {
      for (int j = 0; j < A.p; j++) sX[t * kPredMaxP + j] = 0.0;
    }
    sPx[t] = px;
    sPy[t] = py;
    sY[t] = yy;
    sId[t] = id;
  }
  if (t < A.p) sX0[t] = R.X0[(size_t)t * R.q + i];
  __syncthreads();
  const double2 s0 = R.xy0[i];
  for (int p = t; p < M * (M - 1) / 2; p += kPredBlock) {
    int k = 1;
    while ((k + 1) * k / 2 <= p) k++;  // p = k(k-1)/2 + l, l < k
    const int l = p - k * (k - 1) / 2;
    sD[p] = k < m ? nn_sqrt(sqdist(sPx[k], sPy[k], sPx[l], sPy[l])) : 1e100;
  }
  if (t < M) sD0[t] = t < m ? nn_sqrt(sqdist(sPx[t], sPy[t], s0.x, s0.y)) : 1e100;
  __syncthreads();

  const int s = (int)blockIdx.y * kPredBlock + t;
  if (s >= A.S) return;
  const double *th = A.theta + (size_t)s * A.nTheta;
  const double s2 = th[0], t2 = th[1], phi = th[2];
  const double ip = 1.0 / phi;
  const double phs = phi * A.unit.tsc;  // the unit record's phi is 1
  const bool latent = A.wS != nullptr;
  const double nug = latent ? 0.0 : t2;
  constexpr int T = M * (M + 1) / 2;
  double a[T], v[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    v[k] = s2 * corr_d<COV>(sD0[k], phi, ip, phs, A.unit);
#pragma unroll
    for (int l = 0; l < k; l++)
      a[tri(k, l)] =
          s2 * corr_d<COV>(sD[k * (k - 1) / 2 + l], phi, ip, phs, A.unit);
    a[tri(k, k)] = s2 * 1.0 + nug;
  }
  double vv;
  const bool ok = chol_solve<M>(a, v, vv);  // v <- w
  const double *bs = A.beta + (size_t)s * A.p;
  double x0b = 0.0;
  for (int j = 0; j < A.p; j++) x0b += sX0[j] * bs[j];
  double d = 0.0;
  const double zz = pred_z(R, A.S, s);
  if (latent) {  // padding rows have weight exactly 0 and id 0
    const double *ws = A.wS + (size_t)s * A.n;
#pragma unroll
    for (int k = 0; k < M; k++) d += v[k] * ws[sId[k]];
    const double w0 = sqrt(s2 - vv) * zz + d;
    A.w0out[(size_t)s * R.q + i] = w0;
    R.y0[(size_t)s * R.q + i] =
        sqrt(t2) * A.z2[(size_t)i * A.S + s] + x0b + w0;
  } else {
#pragma unroll
    for (int k = 0; k < M; k++) {
      double xb = 0.0;
      for (int j = 0; j < A.p; j++) xb += sX[k * kPredMaxP + j] * bs[j];
      d += v[k] * (sY[k] - xb);
    }
    R.y0[(size_t)s * R.q + i] = x0b + d + sqrt(s2 + t2 - vv) * zz;
  }
  if (!ok) atomicMin(R.err, i);
}

// Generic path (any m <= 64, any covariance): one wavefront per
// (site, sample), matrix in LDS, one row per lane.
constexpr int kPredCoopWaves = 4;

// Waves per block so that the block's LDS (m(m+1) + 2m doubles per wave)
// stays within 64 KB.
static int pred_coop_waves(int m) {
  const int per = (m * (m + 1) + 2 * m) * (int)sizeof(double);
  const int w = 65536 / per;
  return w < 1 ? 1 : (w > kPredCoopWaves ? kPredCoopWaves : w);
}

template <int MM>
__global__ __launch_bounds__(64 * kPredCoopWaves) void k_pred_coop(
    PredArgs A) {
  extern __shared__ double smem[];
  const int m = A.m, ld = m + 1;
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  double *Am = smem + (size_t)w * (m * ld + 2 * m);
  double *px = Am + m * ld;
  double *py = px + m;
  const SiteRef R = pred_site(A, (int)blockIdx.x);
  const int i = R.i;
  const int s = (int)blockIdx.y * ((int)blockDim.x >> 6) + w;
  if (s >= A.S) return;
  const double *th = A.theta + (size_t)s * A.nTheta;
  const double s2 = th[0], t2 = th[1];
  const CovArgs &c = A.cov_s[s];
  const bool latent = A.wS != nullptr;
  int id = 0;
  if (t < m) {
    id = R.nn0[(size_t)t * R.q + i];
    const double2 pt = A.xy[id];
    px[t] = pt.x;
    py[t] = pt.y;
  }
  wave_sync();
  const double2 s0 = R.xy0[i];
  const int np = m * (m + 1) / 2;
  for (int p = t; p < np; p += 64) {
    int k = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    while ((k + 1) * (k + 2) / 2 <= p) k++;
    while (k * (k + 1) / 2 > p) k--;
    const int l = p - k * (k + 1) / 2;
    Am[k * ld + l] = k == l ? (latent ? c.s2 : c.diag)
                            : s2_corr_rt(sqdist(px[k], py[k], px[l], py[l]), c);
  }
  const double c0 =
      t < m ? s2_corr_rt(sqdist(px[t], py[t], s0.x, s0.y), c) : 0.0;
  wave_sync();
  double x, vv;
  const bool ok = wave_chol_rows<MM>(Am, ld, m, t, c0, x, vv);
  if (!ok) {
    if (t == 0) atomicMin(R.err, i);
    return;
  }
  const double *bs = A.beta + (size_t)s * A.p;
  double part = 0.0;
  if (t < m) {
    if (latent) {
      part = x * A.wS[(size_t)s * A.n + id];
    } else {
      double xb = 0.0;
      for (int j = 0; j < A.p; j++) xb += A.X[(size_t)j * A.n + id] * bs[j];
      part = x * (A.y[id] - xb);
    }
  }
  part = wave_sum(part);
  if (t == 0) {
    double x0b = 0.0;
    for (int j = 0; j < A.p; j++) x0b += R.X0[(size_t)j * R.q + i] * bs[j];
    if (latent) {
      const double w0 = sqrt(s2 - vv) * pred_z(R, A.S, s) + part;
      A.w0out[(size_t)s * R.q + i] = w0;
      R.y0[(size_t)s * R.q + i] =
          sqrt(t2) * A.z2[(size_t)i * A.S + s] + x0b + w0;
    } else {
      R.y0[(size_t)s * R.q + i] =
          x0b + part + sqrt(s2 + t2 - vv) * pred_z(R, A.S, s);
    }
  }
}

int pred_launch(const PredArgs &a, hipStream_t st) {
  if (a.nsite <= 0 || a.S <= 0) return 0;
  if (a.p > kPredMaxP) return -1;
  if (a.cov != 2 && a.m <= kPredMaxM) {
    const dim3 grid(a.nsite, (a.S + kPredBlock - 1) / kPredBlock);
    const dim3 blk(kPredBlock);
#define NNGP_PRED(MM)                                                      \
  do {                                                                    \
    if (a.cov == 0)                                                       \
      hipLaunchKernelGGL((k_pred_lane<MM, 0>), grid, blk, 0, st, a);      \
    else if (a.cov == 1)                                                  \
      hipLaunchKernelGGL((k_pred_lane<MM, 1>), grid, blk, 0, st, a);      \
    else                                                                  \
      hipLaunchKernelGGL((k_pred_lane<MM, 3>), grid, blk, 0, st, a);      \
  } while (0)
    if (a.m <= 4) NNGP_PRED(4);
    else if (a.m <= 8) NNGP_PRED(8);
    else if (a.m <= 10) NNGP_PRED(10);
    else if (a.m <= 12) NNGP_PRED(12);
    else if (a.m <= 15) NNGP_PRED(15);
    else NNGP_PRED(16);
#undef NNGP_PRED
  } else {
    if (a.m > 64) return -1;
    const int nw = pred_coop_waves(a.m);
    const dim3 grid(a.nsite, (a.S + nw - 1) / nw);
    const size_t shm = (size_t)nw * (a.m * (a.m + 1) + 2 * a.m) * sizeof(double);
    switch (wave_rows_for(a.m)) {
      case 16:
        hipLaunchKernelGGL(k_pred_coop<16>, grid, dim3(64 * nw), shm, st, a);
        break;
      case 32:
        hipLaunchKernelGGL(k_pred_coop<32>, grid, dim3(64 * nw), shm, st, a);
        break;
      default:
        hipLaunchKernelGGL(k_pred_coop<64>, grid, dim3(64 * nw), shm, st, a);
    }
  }
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// conjugate-model prediction, one held-out site per lane
// ---------------------------------------------------------------------------
template <int M, int COV>
__global__ __launch_bounds__(kPredBlock) void k_conj_pred_lane(ConjPredArgs A) {
  const int i = (int)blockIdx.x * kPredBlock + (int)threadIdx.x;
  if (i >= A.n0) return;
  const CovArgs &c = A.c;
  int nb[M];
  double px[M], py[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    if (k < A.m) {
      nb[k] = A.nn0[(size_t)k * A.n0 + i];
      const double2 q = A.xy[nb[k]];
      px[k] = q.x;
      py[k] = q.y;
    } else {  // padding: far away, zero correlation, zero coefficient
      nb[k] = 0;
      px[k] = 1e100 * (k + 1);
      py[k] = -1e100 * (k + 1);
    }
  }
  const double2 s0 = A.xy0[i];
  // offsets to the new site in base-2 units (lane_corr)
  constexpr bool kBase2 = lane_base2<COV>();
  if constexpr (kBase2) {
#pragma unroll
    for (int k = 0; k < M; k++) {
      px[k] = (px[k] - s0.x) * c.tsc;
      py[k] = (py[k] - s0.y) * c.tsc;
    }
  }
  constexpr int T = M * (M + 1) / 2;
  double a[T], v[M], c0[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    c0[k] = lane_corr<COV>(kBase2 ? lane_d2<COV>(px[k], py[k])
                                  : sqdist(s0.x, s0.y, px[k], py[k]),
                           c);
    v[k] = c0[k];
#pragma unroll
    for (int l = 0; l < k; l++)
      a[tri(k, l)] = lane_corr<COV>(
          lane_d2<COV>(px[k] - px[l], py[k] - py[l]), c);
    a[tri(k, k)] = c.diag;
  }
  double vv;
  const bool ok = chol_solve<M>(a, v, vv);  // v <- w
  const int p = A.p;
  double yhat = 0.0;
  for (int j = 0; j < p; j++) yhat += A.X0[(size_t)j * A.n0 + i] * A.beta[j];
  double u[kPredMaxP];
  for (int j = 0; j < p; j++) u[j] = 0.0;
  double d = 0.0;
#pragma unroll
  for (int k = 0; k < M; k++) {
    double xb = 0.0;
    for (int j = 0; j < p; j++) {
      const double xkj = A.X[(size_t)j * A.n + nb[k]];
      xb += xkj * A.beta[j];
      u[j] = fma(xkj, v[k], u[j]);
    }
    d += v[k] * (A.y[nb[k]] - xb);
  }
  yhat += d;
  for (int j = 0; j < p; j++) u[j] = A.X0[(size_t)j * A.n0 + i] - u[j];
  // u^T Vb u with Vb symmetric, lower triangle stored
  double q = 0.0;
  for (int j = 0; j < p; j++) {
    double r = 0.0;
    for (int k = 0; k < p; k++) {
      const double vjk = j >= k ? A.Vb[k * p + j] : A.Vb[j * p + k];
      r += vjk * u[k];
    }
    q += u[j] * r;
  }
  A.y0[i] = yhat;
  A.y0var[i] = A.b * (q + 1.0 + A.alpha - vv) / (A.a - 1.0);
  if (!ok) atomicMin(A.err, i);
}

template <int MM>
__global__ __launch_bounds__(64 * kPredCoopWaves) void k_conj_pred_coop(
    ConjPredArgs A) {
  extern __shared__ double smem[];
  const int m = A.m, ld = m + 1;
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  double *Am = smem + (size_t)w * (m * ld + 2 * m);
  double *px = Am + m * ld;
  double *py = px + m;
  const int i = (int)blockIdx.x * ((int)blockDim.x >> 6) + w;
  if (i >= A.n0) return;
  const CovArgs &c = A.c;
  int id = 0;
  if (t < m) {
    id = A.nn0[(size_t)t * A.n0 + i];
    const double2 pt = A.xy[id];
    px[t] = pt.x;
    py[t] = pt.y;
  }
  wave_sync();
  const double2 s0 = A.xy0[i];
  const int np = m * (m + 1) / 2;
  for (int p = t; p < np; p += 64) {
    int k = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    while ((k + 1) * (k + 2) / 2 <= p) k++;
    while (k * (k + 1) / 2 > p) k--;
    const int l = p - k * (k + 1) / 2;
    Am[k * ld + l] = k == l ? c.diag
                            : s2_corr_rt(sqdist(px[k], py[k], px[l], py[l]), c);
  }
  const double c0 =
      t < m ? s2_corr_rt(sqdist(px[t], py[t], s0.x, s0.y), c) : 0.0;
  wave_sync();
  double x, vv;
  if (!wave_chol_rows<MM>(Am, ld, m, t, c0, x, vv)) {
    if (t == 0) atomicMin(A.err, i);
    return;
  }
  const int p = A.p;
  double part = 0.0;
  double u[kPredMaxP];
  for (int j = 0; j < p; j++) {
    const double xj = t < m ? A.X[(size_t)j * A.n + id] : 0.0;
    u[j] = wave_sum(x * xj);
  }
  if (t < m) {
    double xb = 0.0;
    for (int j = 0; j < p; j++) xb += A.X[(size_t)j * A.n + id] * A.beta[j];
    part = x * (A.y[id] - xb);
  }
  part = wave_sum(part);
  if (t == 0) {
    double yhat = 0.0;
    for (int j = 0; j < p; j++) yhat += A.X0[(size_t)j * A.n0 + i] * A.beta[j];
    for (int j = 0; j < p; j++) u[j] = A.X0[(size_t)j * A.n0 + i] - u[j];
    double q = 0.0;
    for (int j = 0; j < p; j++) {
      double r = 0.0;
      for (int k = 0; k < p; k++) {
        const double vjk = j >= k ? A.Vb[k * p + j] : A.Vb[j * p + k];
        r += vjk * u[k];
      }
      q += u[j] * r;
    }
    A.y0[i] = yhat + part;
    A.y0var[i] = A.b * (q + 1.0 + A.alpha - vv) / (A.a - 1.0);
  }
}

int conj_pred_launch(const ConjPredArgs &a, hipStream_t st) {
  if (a.n0 <= 0) return 0;
  if (a.p > kPredMaxP) return -1;
  if (a.c.cov != 2 && a.m <= kPredMaxM) {
    const dim3 grid((a.n0 + kPredBlock - 1) / kPredBlock), blk(kPredBlock);
#define NNGP_CPRED(MM)                                                       \
  do {                                                                      \
    if (a.c.cov == 0)                                                       \
      hipLaunchKernelGGL((k_conj_pred_lane<MM, 0>), grid, blk, 0, st, a);   \
    else if (a.c.cov == 1)                                                  \
      hipLaunchKernelGGL((k_conj_pred_lane<MM, 1>), grid, blk, 0, st, a);   \
    else                                                                    \
      hipLaunchKernelGGL((k_conj_pred_lane<MM, 3>), grid, blk, 0, st, a);   \
  } while (0)
    if (a.m <= 4) NNGP_CPRED(4);
    else if (a.m <= 8) NNGP_CPRED(8);
    else if (a.m <= 10) NNGP_CPRED(10);
    else if (a.m <= 12) NNGP_CPRED(12);
    else if (a.m <= 15) NNGP_CPRED(15);
    else NNGP_CPRED(16);
#undef NNGP_CPRED
  } else {
    if (a.m > 64) return -1;
    const int nw = pred_coop_waves(a.m);
    const dim3 grid((a.n0 + nw - 1) / nw);
    const size_t shm = (size_t)nw * (a.m * (a.m + 1) + 2 * a.m) * sizeof(double);
    switch (wave_rows_for(a.m)) {
      case 16:
        hipLaunchKernelGGL(k_conj_pred_coop<16>, grid, dim3(64 * nw), shm, st,
                           a);
        break;
      case 32:
        hipLaunchKernelGGL(k_conj_pred_coop<32>, grid, dim3(64 * nw), shm, st,
                           a);
        break;
      default:
        hipLaunchKernelGGL(k_conj_pred_coop<64>, grid, dim3(64 * nw), shm, st,
                           a);
    }
  }
  return (int)hipGetLastError();
}

}  // namespace nngp
