// nngp_cvgrid.hip — spConjNNGP's k-fold search over a (phi, alpha[, nu])
// grid (R/spConjNNGP.R:287-357, cNNGP of src/cNNGP.cpp:73-358 per fold) as
// one captured HIP graph.
//
// The k folds are fixed at creation: each fold's training set (the sites
// not held out, in their NNGP order), its neighbor index (mkNNIndxCB on the
// training set) and its held-out sites with their nn2 neighbors stay in HBM
// at fixed strides.  A run over g grid points is k x g problems, done by
// five batched launches recorded once into a graph:
//   1. updateConjBF of every problem with the cross-products
//      X'V^-1X, X'V^-1y, y'V^-1y summed per block (k_bf_lane / k_bf_coop,
//      BATCH), launch shape chosen by the autotuner;
//   2. the column sums of every problem;
//   3. per problem beta.hat = (X'V^-1X)^-1 X'V^-1y, a and b
//      (src/cNNGP.cpp:227-245);
//   4. the prediction of every held-out site of every problem
//      (src/cNNGP.cpp:255-299) and its CRPS and squared-error terms
//      (R/util.R:247-255), summed per block;
//   5. the column sums of those terms.
// The host then forms the scores of R/spConjNNGP.R:336-350: per grid point
// the mean over folds of the per-fold mean CRPS and of the per-fold RMSPE.
#include <climits>
#include <cstring>

#include "nngp_host.hpp"
#include "nngp_strips.hpp"
#include "nngp_wave.hpp"

namespace nngp {

constexpr int kSol = 96;        // per-problem solution record (doubles)
constexpr int kSolA = 8, kSolB = 9, kSolOk = 10, kSolH = 16, kSolV = 32;
constexpr int kCvPredThreads = 256;
constexpr int kCvCoopWaves = 4;
constexpr int kCvCoopBlocks = 512;

__global__ void k_fill_ints(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// per-problem solve: sums -> beta.hat, a, b, (X'V^-1X)^-1
// ---------------------------------------------------------------------------
template <int P>
__global__ void k_conj_solve(const double *sums, int nprob, double a0,
                             double b0, int n, double *sol) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nprob) return;
  const double *raw = sums + (size_t)j * kStatsMax;
  // gram entry (r, s), r >= s: raw[2 + r(r+1)/2 + s]; X columns 0..P-1, y P
  double L[P][P], h[P];
#pragma unroll
  for (int r = 0; r < P; r++) {
#pragma unroll
    for (int s = 0; s <= r; s++) L[r][s] = raw[2 + r * (r + 1) / 2 + s];
    h[r] = raw[2 + P * (P + 1) / 2 + r];
  }
  const double yy = raw[2 + P * (P + 1) / 2 + P];
  bool ok = true;
#pragma unroll
  for (int c = 0; c < P; c++) {
    double d = L[c][c];
#pragma unroll
    for (int k = 0; k < c; k++) d -= L[c][k] * L[c][k];
    ok = ok && d > 0.0;
    d = sqrt(d > 0.0 ? d : 1.0);
    L[c][c] = d;
#pragma unroll
    for (int r = c + 1; r < P; r++) {
      double s = L[r][c];
#pragma unroll
      for (int k = 0; k < c; k++) s -= L[r][k] * L[c][k];
      L[r][c] = s / d;
    }
  }
  // Linv = L^-1 (lower), then G^-1 = Linv^T Linv
  double Li[P][P];
#pragma unroll
  for (int c = 0; c < P; c++) {
#pragma unroll
    for (int r = 0; r < P; r++) Li[r][c] = 0.0;
    Li[c][c] = 1.0 / L[c][c];
#pragma unroll
    for (int r = c + 1; r < P; r++) {
      double s = 0.0;
#pragma unroll
      for (int k = c; k < r; k++) s += L[r][k] * Li[k][c];
      Li[r][c] = -s / L[r][r];
    }
  }
  double *o = sol + (size_t)j * kSol;
  double beta[P];
#pragma unroll
  for (int r = 0; r < P; r++) {
    double br = 0.0;
#pragma unroll
    for (int s = 0; s < P; s++) {
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < P; k++) {
        if (k >= r && k >= s) v += Li[k][r] * Li[k][s];
      }
      o[kSolV + s * P + r] = v;  // column-major p x p, full
      br += v * h[s];
    }
    beta[r] = br;
    o[r] = br;
    o[kSolH + r] = h[r];
  }
  double bh = 0.0;
#pragma unroll
  for (int r = 0; r < P; r++) bh += beta[r] * h[r];
  o[kSolA] = a0 + 1.0 * n / 2.0;
  o[kSolB] = b0 + 0.5 * (yy - bh);
  o[kSolOk] = ok ? 1.0 : 0.0;
}

static int solve_launch(const double *sums, int nprob, int p, double a0,
                        double b0, int n, double *sol, hipStream_t st) {
  const dim3 grid((nprob + 63) / 64), blk(64);
  switch (p) {
#define NNGP_SOLVE(P)                                                       \
  case P:                                                                  \
    hipLaunchKernelGGL(k_conj_solve<P>, grid, blk, 0, st, sums, nprob, a0, \
                       b0, n, sol);                                         \
    break;
    NNGP_SOLVE(1) NNGP_SOLVE(2) NNGP_SOLVE(3) NNGP_SOLVE(4)
    NNGP_SOLVE(5) NNGP_SOLVE(6) NNGP_SOLVE(7) NNGP_SOLVE(8)
#undef NNGP_SOLVE
    default: return -1;
  }
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// held-out prediction and scoring, all problems in one launch
// ---------------------------------------------------------------------------
struct CvPredArgs {
  const double2 *xy;     // training coordinates, fold f at xy + f*nmod
  const double *Z;       // training X|y rows (zs doubles), fold f at
  int nmod, zs, p, m;    //   Z + f*nmod*zs
  const double2 *xy0;    // held-out coordinates, fold f at xy0 + f*nho
  const double *X0;      // held-out X (nho x p column-major per fold)
  const double *y0;      // held-out y
  const int *nn0;        // held-out neighbors (nho x m column-major per fold)
  int nho, ng;
  const CovArgs *covs;   // per grid point
  const double *sol;     // per problem
  double *prow;          // per problem: 2 columns x prow_cap partial rows
  int prow_cap;
  int coop_blocks;
  double *y0hat, *y0var; // optional: nho per problem
  int *perr;             // per problem
};

__device__ __forceinline__ void crps_sq(double yobs, double yhat, double var,
                                        double &crps, double &sq) {
  // R/util.R:247-251 with dnorm and pnorm written out
  const double sd = sqrt(var);
  const double z = (yobs - yhat) / sd;
  const double dn = 0.3989422804014327 * exp(-0.5 * z * z);
  const double pn = 0.5 * erfc(-z * 0.7071067811865476);
  crps = -sd * (0.5641895835477563 - 2.0 * dn - z * (2.0 * pn - 1.0));
  sq = (yobs - yhat) * (yobs - yhat);
}

// u^T Vb u with Vb p x p column-major, full.
__device__ __forceinline__ double quad_vb(const double *Vb, const double *u,
                                          int p) {
  double q = 0.0;
  for (int r = 0; r < p; r++) {
    double s = 0.0;
    for (int c = 0; c < p; c++) s += Vb[c * p + r] * u[c];
    q += u[r] * s;
  }
  return q;
}

__device__ inline void block_pair_sum(double a, double b, double *prow,
                                      int cap, int row) {
  __shared__ double red[kCvPredThreads / 64][2];
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  const int nw = (int)blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  if (t == 0) {
    red[w][0] = a;
    red[w][1] = b;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int ww = 0; ww < nw; ww++) s += red[ww][threadIdx.x];
    prow[(size_t)threadIdx.x * cap + row] = s;
  }
}

// One held-out site per lane (closed-form covariances, m <= 16).
template <int M, int COV>
__global__ __launch_bounds__(kCvPredThreads) void k_cv_pred_lane(
    CvPredArgs A) {
  const int nb = (A.nho + kCvPredThreads - 1) / kCvPredThreads;
  const int b = xcd_block(nb);
  if (b < 0) return;
  const int j = (int)blockIdx.y;
  const int f = j / A.ng;
  const CovArgs c = A.covs[j % A.ng];
  const double *so = A.sol + (size_t)j * kSol;
  const int i = b * kCvPredThreads + (int)threadIdx.x;
  double crps = 0.0, sq = 0.0;
  if (i < A.nho) {
    const double2 *xy = A.xy + (size_t)f * A.nmod;
    const double *Z = A.Z + (size_t)f * A.nmod * A.zs;
    const int *nn0 = A.nn0 + (size_t)f * A.nho * A.m;
    const double *X0 = A.X0 + (size_t)f * A.nho * A.p;
    int nbr[M];
    double px[M], py[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k < A.m) {
        nbr[k] = nn0[(size_t)k * A.nho + i];
        const double2 q = xy[nbr[k]];
        px[k] = q.x;
        py[k] = q.y;
      } else {  // padding: far away, zero correlation, zero coefficient
        nbr[k] = 0;
        px[k] = 1e100 * (k + 1);
        py[k] = -1e100 * (k + 1);
      }
    }
    const double2 s0 = A.xy0[(size_t)f * A.nho + i];
    // offsets to the held-out site in base-2 units (lane_corr)
    constexpr bool kBase2 = lane_base2<COV>();
    if constexpr (kBase2) {
#pragma unroll
      for (int k = 0; k < M; k++) {
        px[k] = (px[k] - s0.x) * c.tsc;
        py[k] = (py[k] - s0.y) * c.tsc;
      }
    }
    constexpr int T = M * (M + 1) / 2;
    double a[T], v[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
      v[k] = lane_corr<COV>(kBase2 ? lane_d2<COV>(px[k], py[k])
                                   : sqdist(s0.x, s0.y, px[k], py[k]),
                            c);
#pragma unroll
      for (int l = 0; l < k; l++)
        a[tri(k, l)] = lane_corr<COV>(
          lane_d2<COV>(px[k] - px[l], py[k] - py[l]), c);
      a[tri(k, k)] = c.diag;
    }
    double vv;
    const bool ok = chol_solve<M>(a, v, vv);  // v <- w
    const int p = A.p;
    double u[kMaxP];
    double yhat = 0.0;
    for (int r = 0; r < p; r++) {
      u[r] = 0.0;
      yhat += X0[(size_t)r * A.nho + i] * so[r];
    }
    double d = 0.0;
#pragma unroll
    for (int k = 0; k < M; k++) {
      const double *zr = Z + (size_t)nbr[k] * A.zs;
      double xb = 0.0;
      for (int r = 0; r < p; r++) {
        xb += zr[r] * so[r];
        u[r] = fma(zr[r], v[k], u[r]);
      }
      d += v[k] * (zr[p] - xb);
    }
    yhat += d;
    for (int r = 0; r < p; r++) u[r] = X0[(size_t)r * A.nho + i] - u[r];
    const double var = so[kSolB] *
                       (quad_vb(so + kSolV, u, p) + 1.0 + c.t2 - vv) /
                       (so[kSolA] - 1.0);
    if (A.y0hat) {
      A.y0hat[(size_t)j * A.nho + i] = yhat;
      A.y0var[(size_t)j * A.nho + i] = var;
    }
    crps_sq(A.y0[(size_t)f * A.nho + i], yhat, var, crps, sq);
    if (!ok) atomicMin(A.perr + j, i);
  }
  block_pair_sum(crps, sq, A.prow + (size_t)j * 2 * A.prow_cap, A.prow_cap,
                 b);
}

// One held-out site per wavefront (any m <= MM <= 64, any covariance);
// the waves of one XCD's logical blocks take its run of sites round robin
// (xcd_site_run).
template <int MM>
__global__ __launch_bounds__(64 * kCvCoopWaves) void k_cv_pred_coop(
    CvPredArgs A) {
  extern __shared__ double smem[];
  const int nw = (int)blockDim.x >> 6;
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  const int b = xcd_block(A.coop_blocks);
  if (b < 0) return;
  const int j = (int)blockIdx.y;
  const int f = j / A.ng;
  const CovArgs c = A.covs[j % A.ng];
  const double *so = A.sol + (size_t)j * kSol;
  const int m = A.m, ld = m + 1, p = A.p;
  double *Am = smem + (size_t)w * (m * ld + 2 * m);
  double *px = Am + m * ld;
  double *py = px + m;
  const double2 *xy = A.xy + (size_t)f * A.nmod;
  const double *Z = A.Z + (size_t)f * A.nmod * A.zs;
  const int *nn0 = A.nn0 + (size_t)f * A.nho * m;
  const double *X0 = A.X0 + (size_t)f * A.nho * p;
  const SiteRun run = xcd_site_run(A.coop_blocks, b, A.nho, nw, w);
  double crps_acc = 0.0, sq_acc = 0.0;
  for (int i = run.first; i < run.end; i += run.stride) {
    int id = 0;
    if (t < m) {
      id = nn0[(size_t)t * A.nho + i];
      const double2 q = xy[id];
      px[t] = q.x;
      py[t] = q.y;
    }
    wave_sync();
    const double2 s0 = A.xy0[(size_t)f * A.nho + i];
    const int np = m * (m + 1) / 2;
    for (int q = t; q < np; q += 64) {
      int k = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
      while ((k + 1) * (k + 2) / 2 <= q) k++;
      while (k * (k + 1) / 2 > q) k--;
      const int l = q - k * (k + 1) / 2;
      Am[k * ld + l] =
          k == l ? c.diag : s2_corr_rt(sqdist(px[k], py[k], px[l], py[l]), c);
    }
    const double c0 = t < m ? s2_corr_rt(sqdist(px[t], py[t], s0.x, s0.y), c)
                            : 0.0;
    wave_sync();
    double x, vv;
    if (!wave_chol_rows<MM>(Am, ld, m, t, c0, x, vv)) {
      if (t == 0) atomicMin(A.perr + j, i);
      wave_sync();
      continue;
    }
    const double *zr = Z + (size_t)id * A.zs;
    double u[kMaxP];
    double part = 0.0;
    {
      double xb = 0.0;
      for (int r = 0; r < p; r++) {
        const double xr = t < m ? zr[r] : 0.0;
        u[r] = wave_sum(x * xr);
        xb += xr * so[r];
      }
      part = wave_sum(t < m ? x * (zr[p] - xb) : 0.0);
    }
    double yhat = part;
    for (int r = 0; r < p; r++) {
      const double x0r = X0[(size_t)r * A.nho + i];
      yhat += x0r * so[r];
      u[r] = x0r - u[r];
    }
    const double var = so[kSolB] *
                       (quad_vb(so + kSolV, u, p) + 1.0 + c.t2 - vv) /
                       (so[kSolA] - 1.0);
    if (t == 0) {
      if (A.y0hat) {
        A.y0hat[(size_t)j * A.nho + i] = yhat;
        A.y0var[(size_t)j * A.nho + i] = var;
      }
      double cr, s2;
      crps_sq(A.y0[(size_t)f * A.nho + i], yhat, var, cr, s2);
      crps_acc += cr;
      sq_acc += s2;
    }
    wave_sync();  // this LDS area is reused by the wave's next site
  }
  block_pair_sum(crps_acc, sq_acc, A.prow + (size_t)j * 2 * A.prow_cap,
                 A.prow_cap, b);
}

// Rows (logical blocks) per problem of the prediction launch.
static int cv_pred_rows(int nho, bool lane) {
  if (lane) return (nho + kCvPredThreads - 1) / kCvPredThreads;
  const int need = (nho + kCvCoopWaves - 1) / kCvCoopWaves;
  return need < kCvCoopBlocks ? need : kCvCoopBlocks;
}

static int cv_pred_launch(CvPredArgs a, int variant, int nprob,
                          hipStream_t st) {
  const bool lane = a.m <= 16 && variant >= 0;
  const int nrows = cv_pred_rows(a.nho, lane);
  const dim3 grid(8 * ((nrows + 7) / 8), nprob);
  if (lane) {
    const dim3 blk(kCvPredThreads);
#define NNGP_CVP_COV(MM)                                                           \
  do {                                                                            \
    switch (variant) {                                                            \
      case 0: hipLaunchKernelGGL((k_cv_pred_lane<MM, 0>), grid, blk, 0, st, a); break; \
      case 1: hipLaunchKernelGGL((k_cv_pred_lane<MM, 1>), grid, blk, 0, st, a); break; \
      case kCovMatern15: hipLaunchKernelGGL((k_cv_pred_lane<MM, kCovMatern15>), grid, blk, 0, st, a); break; \
      case kCovMatern25: hipLaunchKernelGGL((k_cv_pred_lane<MM, kCovMatern25>), grid, blk, 0, st, a); break; \
      default: hipLaunchKernelGGL((k_cv_pred_lane<MM, 3>), grid, blk, 0, st, a); break; \
    }                                                                             \
  } while (0)
    if (a.m <= 4) NNGP_CVP_COV(4);
    else if (a.m <= 8) NNGP_CVP_COV(8);
    else if (a.m <= 10) NNGP_CVP_COV(10);
    else if (a.m <= 12) NNGP_CVP_COV(12);
    else if (a.m <= 15) NNGP_CVP_COV(15);
    else NNGP_CVP_COV(16);
#undef NNGP_CVP_COV
  } else {
    a.coop_blocks = nrows;
    int waves = kCvCoopWaves;
    const size_t per = (size_t)(a.m * (a.m + 1) + 2 * a.m) * sizeof(double);
    while (waves > 1 && waves * per > 65536) waves--;
    // fewer waves per block only lengthen each wave's walk over its XCD's
    // run of sites
    const dim3 blk(64 * waves);
    switch (wave_rows_for(a.m)) {
      case 16:
        hipLaunchKernelGGL(k_cv_pred_coop<16>, grid, blk, waves * per, st, a);
        break;
      case 32:
        hipLaunchKernelGGL(k_cv_pred_coop<32>, grid, blk, waves * per, st, a);
        break;
      default:
        hipLaunchKernelGGL(k_cv_pred_coop<64>, grid, blk, waves * per, st, a);
    }
  }
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// the grid context
// ---------------------------------------------------------------------------
struct CvGrid {
  int n = 0, p = 0, m = 0, k = 0, nho = 0, nmod = 0, cov = 0, nz = 0, zs = 0;
  int search = 0;  // per-fold neighbor search: 0 code book, 1 scan
  long long nIndx = 0, nn_stride = 0;
  double a0 = 0.0, b0 = 0.0;
  Stream st;
  DevBuf xy, nnI, lu, Z, xy0, X0, y0, nn0;
  // run state, sized for cap_prob problems
  int cap_prob = 0, rows_cap = 0, prow_cap = 0;
  DevBuf covs, rows, sums, errs, errd, sol, prow, pscore, perr, perrd;
  DevBuf y0hat, y0var;
  PinnedBuf hbuf;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_g = -1, graph_variant = -99;
  bool graph_y0 = false;
  BfConfig cfg;
  BfConfig pinned{0, 0, 0};  // launch shape from nngp_opts.shape (threads 0:
                             // none; the autotuner chooses)
  BfTuneReport tune;
  bool tuned = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  int create(const double *yv, const double *X, int p_, int n_,
             const double *coords, int m_, int k_, const int *ho, int nho_,
             const double *ig, int cov_, int search_);
  int run(const double *thetaAlpha, int g, double *scores,
          nngp_cvgrid_detail *det);
  int record(int g, int variant, bool want_y0);
  ~CvGrid() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

int CvGrid::create(const double *yv, const double *X, int p_, int n_,
                   const double *coords, int m_, int k_, const int *ho,
                   int nho_, const double *ig, int cov_, int search_) {
  n = n_;
  search = search_;
  p = p_;
  m = m_;
  k = k_;
  nho = nho_;
  nmod = n - nho;
  cov = cov_;
  nz = p + 1;
  zs = (nz + 1) & ~1;
  a0 = ig[0];
  b0 = ig[1];
  nIndx = n_indx(nmod, m);
  nn_stride = nIndx + 64;
  NNGP_TRY(st.create());
  NNGP_HIP(hipEventCreate(&ev0));
  NNGP_HIP(hipEventCreate(&ev1));
  NNGP_TRY(xy.alloc(sizeof(double2) * (size_t)k * nmod));
  NNGP_TRY(nnI.alloc(sizeof(int) * (size_t)k * nn_stride));
  NNGP_TRY(lu.alloc(sizeof(int) * 2 * (size_t)k * nmod));
  NNGP_TRY(Z.alloc(sizeof(double) * (size_t)k * nmod * zs));
  NNGP_TRY(xy0.alloc(sizeof(double2) * (size_t)k * nho));
  NNGP_TRY(X0.alloc(sizeof(double) * (size_t)k * nho * p));
  NNGP_TRY(y0.alloc(sizeof(double) * (size_t)k * nho));
  NNGP_TRY(nn0.alloc(sizeof(int) * (size_t)k * nho * m));
  NNGP_HIP(hipMemsetAsync(nnI.p, 0, nnI.bytes, st.s));
  std::vector<char> held(n);
  std::vector<double> cm((size_t)2 * nmod), c0((size_t)2 * nho);
  std::vector<double> hz((size_t)nmod * zs, 0.0), hx0((size_t)nho * p),
      hy0(nho);
  for (int f = 0; f < k; f++) {
    const int *hf = ho + (size_t)f * nho;
    std::fill(held.begin(), held.end(), 0);
    for (int t = 0; t < nho; t++) {
      if (hf[t] < 0 || hf[t] >= n || held[hf[t]])
        return set_err(NNGP_ERR_ARG,
                       "fold %d: held-out index %d repeated or outside [0, n)",
                       f, hf[t]);
      held[hf[t]] = 1;
    }
    int r = 0;
    for (int i = 0; i < n; i++) {
      if (held[i]) continue;
      cm[r] = coords[i];
      cm[(size_t)nmod + r] = coords[(size_t)n + i];
      for (int j = 0; j < p; j++) hz[(size_t)r * zs + j] = X[(size_t)j * n + i];
      hz[(size_t)r * zs + p] = yv[i];
      r++;
    }
    for (int s = 0; s < nho; s++) {  // held-out sites in the given order
      const int i = hf[s];
      c0[s] = coords[i];
      c0[(size_t)nho + s] = coords[(size_t)n + i];
      for (int j = 0; j < p; j++) hx0[(size_t)j * nho + s] = X[(size_t)j * n + i];
      hy0[s] = yv[i];
    }
    NNGP_HIP(hipMemcpyAsync(Z.as<double>() + (size_t)f * nmod * zs, hz.data(),
                            sizeof(double) * hz.size(), hipMemcpyHostToDevice,
                            st.s));
    NNGP_HIP(hipMemcpyAsync(X0.as<double>() + (size_t)f * nho * p, hx0.data(),
                            sizeof(double) * hx0.size(), hipMemcpyHostToDevice,
                            st.s));
    NNGP_HIP(hipMemcpyAsync(y0.as<double>() + (size_t)f * nho, hy0.data(),
                            sizeof(double) * nho, hipMemcpyHostToDevice, st.s));
    // the fold's training coordinates and their neighbor index
    DevBuf dxy, dxy0;
    NNGP_TRY(upload_xy(cm.data(), nmod, dxy, st.s));
    NNGP_HIP(hipMemcpyAsync(xy.as<double2>() + (size_t)f * nmod, dxy.p,
                            sizeof(double2) * nmod, hipMemcpyDeviceToDevice,
                            st.s));
    HostStrips hs;
    build_strips(nmod, cm.data(), cm.data() + nmod, m, hs);
    DevStrips ds;
    NNGP_TRY(ds.upload(hs, st.s));
    NNGP_TRY(knn_build(cm.data(), nmod, m, search, hs, ds, dxy.as<double2>(),
                       nnI.as<int>() + (size_t)f * nn_stride, nullptr,
                       lu.as<int>() + 2 * (size_t)f * nmod, st.s, nullptr));
    // held-out sites and their m nearest training sites
    NNGP_TRY(upload_xy(c0.data(), nho, dxy0, st.s));
    NNGP_HIP(hipMemcpyAsync(xy0.as<double2>() + (size_t)f * nho, dxy0.p,
                            sizeof(double2) * nho, hipMemcpyDeviceToDevice,
                            st.s));
    Nn2Args na;  // same strips: nn2 drops the predecessor constraint

    na.s = ds.idx;
    na.q = dxy0.as<double2>();
    na.nq = nho;
    na.k = m;
    na.idx = nn0.as<int>() + (size_t)f * nho * m;
    na.dist = nullptr;
    if (nn2_launch(na, st.s)) return set_err(NNGP_ERR_HIP, "nn2 launch failed");
    NNGP_HIP(hipStreamSynchronize(st.s));  // before the staging is reused
  }
  return NNGP_OK;
}

// Record the run's launches for g grid points into the graph.
int CvGrid::record(int g, int variant, bool want_y0) {
  const int nprob = k * g;
  if (exec) {
    (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  if (graph) {
    (void)hipGraphDestroy(graph);
    graph = nullptr;
  }
  BfArgs a;
  a.xy = xy.as<double2>();
  a.nnIndx = nnI.as<int>();
  a.lu = lu.as<int>();
  a.n = nmod;
  a.m = m;
  a.B = nullptr;
  a.F = nullptr;
  a.Z = Z.as<double>();
  a.nz = nz;
  a.zs = zs;
  a.err = errs.as<int>();
  a.rows = rows.as<double>();
  a.rows_cap = rows_cap;
  std::memset(&a.c, 0, sizeof a.c);
  a.variant = variant;
  a.ng = g;
  a.covs = covs.as<CovArgs>();
  a.fs_xy = nmod;
  a.fs_nn = nn_stride;
  a.fs_lu = 2LL * nmod;
  a.fs_Z = (long long)nmod * zs;
  a.ps_rows = (long long)rows_cap * (2 + nz * (nz + 1) / 2);
  CvPredArgs pa;
  pa.xy = xy.as<double2>();
  pa.Z = Z.as<double>();
  pa.nmod = nmod;
  pa.zs = zs;
  pa.p = p;
  pa.m = m;
  pa.xy0 = xy0.as<double2>();
  pa.X0 = X0.as<double>();
  pa.y0 = y0.as<double>();
  pa.nn0 = nn0.as<int>();
  pa.nho = nho;
  pa.ng = g;
  pa.covs = covs.as<CovArgs>();
  pa.sol = sol.as<double>();
  pa.prow = prow.as<double>();
  pa.prow_cap = prow_cap;
  pa.coop_blocks = 0;
  pa.y0hat = want_y0 ? y0hat.as<double>() : nullptr;
  pa.y0var = want_y0 ? y0var.as<double>() : nullptr;
  pa.perr = perr.as<int>();
  const bool lane_pred = m <= 16 && variant >= 0;
  NNGP_HIP(hipStreamBeginCapture(st.s, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  const dim3 fb((nprob + 255) / 256), ft(256);
  hipLaunchKernelGGL(k_fill_ints, fb, ft, 0, st.s, errs.as<int>(), nprob,
                     INT_MAX);
  hipLaunchKernelGGL(k_fill_ints, fb, ft, 0, st.s, perr.as<int>(), nprob,
                     INT_MAX);
  if (!rc) rc = bf_launch(a, cfg, nprob, st.s);
  if (!rc)
    rc = bf_sums_launch(a, cfg, nprob, sums.as<double>(), kStatsMax,
                        errd.as<double>(), st.s);
  if (!rc)
    rc = solve_launch(sums.as<double>(), nprob, p, a0, b0, nmod,
                      sol.as<double>(), st.s);
  if (!rc) rc = cv_pred_launch(pa, variant, nprob, st.s);
  if (!rc)
    rc = rows_sum_launch(prow.as<double>(), prow_cap,
                         cv_pred_rows(nho, lane_pred), 2LL * prow_cap, 2,
                         nprob, pscore.as<double>(), 2, perr.as<int>(),
                         perrd.as<double>(), st.s);
  const hipError_t e = hipStreamEndCapture(st.s, &graph);
  if (rc) return set_err(NNGP_ERR_HIP, "grid launch capture failed (%d)", rc);
  NNGP_HIP(e);
  NNGP_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  graph_g = g;
  graph_variant = variant;
  graph_y0 = want_y0;
  return NNGP_OK;
}

int CvGrid::run(const double *thetaAlpha, int g, double *scores,
                nngp_cvgrid_detail *det) {
  const int nTheta = cov == 2 ? 3 : 2;
  const int nprob = k * g;
  // one launch variant for the whole grid: the templated closed form when
  // every point has the same matern smoothness with one, else the generic
  // (runtime) evaluation
  std::vector<CovArgs> hc(g);
  int variant = cov_variant(cov, cov == 2 ? thetaAlpha[2] : 0.0);
  for (int j = 0; j < g; j++) {
    const double *t = thetaAlpha + (size_t)j * nTheta;
    const double nu = cov == 2 ? t[2] : 0.0;
    if (!(t[0] > 0.0) || !(t[1] >= 0.0) || (cov == 2 && !(nu > 0.0)))
      return set_err(NNGP_ERR_ARG, "grid point %d: need phi > 0, alpha >= 0%s",
                     j, cov == 2 ? ", nu > 0" : "");
    hc[j] = make_cov(cov, 1.0, t[1], t[0], nu);
    if (cov_variant(cov, nu) != variant) variant = -1;
  }
  const bool want_y0 = det && (det->y0Hat || det->y0HatVar);
  const int nv = 2 + nz * (nz + 1) / 2;
  if (nprob > cap_prob) {
    rows_cap = rows_capacity(nmod);
    prow_cap = cv_pred_rows(nho, true) > cv_pred_rows(nho, false)
                   ? cv_pred_rows(nho, true)
                   : cv_pred_rows(nho, false);
    NNGP_TRY(covs.alloc(sizeof(CovArgs) * (size_t)g));
    NNGP_TRY(rows.alloc(sizeof(double) * (size_t)nprob * nv * rows_cap));
    NNGP_TRY(sums.alloc(sizeof(double) * (size_t)nprob * kStatsMax));
    NNGP_TRY(errs.alloc(sizeof(int) * (size_t)nprob));
    NNGP_TRY(errd.alloc(sizeof(double) * (size_t)nprob));
    NNGP_TRY(sol.alloc(sizeof(double) * (size_t)nprob * kSol));
    NNGP_TRY(prow.alloc(sizeof(double) * (size_t)nprob * 2 * prow_cap));
    NNGP_TRY(pscore.alloc(sizeof(double) * (size_t)nprob * 2));
    NNGP_TRY(perr.alloc(sizeof(int) * (size_t)nprob));
    NNGP_TRY(perrd.alloc(sizeof(double) * (size_t)nprob));
    NNGP_TRY(hbuf.alloc(sizeof(double) * (size_t)nprob * (kSol + 4)));
    cap_prob = nprob;
    graph_g = -1;
  } else if (covs.bytes < sizeof(CovArgs) * (size_t)g) {
    NNGP_TRY(covs.alloc(sizeof(CovArgs) * (size_t)g));
    graph_g = -1;
  }
  if (want_y0 && y0hat.bytes < sizeof(double) * (size_t)nprob * nho) {
    NNGP_TRY(y0hat.alloc(sizeof(double) * (size_t)nprob * nho));
    NNGP_TRY(y0var.alloc(sizeof(double) * (size_t)nprob * nho));
    graph_g = -1;
  }
  NNGP_HIP(hipMemcpyAsync(covs.p, hc.data(), sizeof(CovArgs) * (size_t)g,
                          hipMemcpyHostToDevice, st.s));
  if (!tuned || graph_variant != variant) {
    // autotune the updateConjBF launch on the real problems (a bounded
    // number of them), then keep the choice for this context
    BfArgs a;
    a.xy = xy.as<double2>();
    a.nnIndx = nnI.as<int>();
    a.lu = lu.as<int>();
    a.n = nmod;
    a.m = m;
    a.B = nullptr;
    a.F = nullptr;
    a.Z = Z.as<double>();
    a.nz = nz;
    a.zs = zs;
    a.err = errs.as<int>();
    a.rows = rows.as<double>();
    a.rows_cap = rows_cap;
    std::memset(&a.c, 0, sizeof a.c);
    a.variant = variant;
    a.ng = g;
    a.covs = covs.as<CovArgs>();
    a.fs_xy = nmod;
    a.fs_nn = nn_stride;
    a.fs_lu = 2LL * nmod;
    a.fs_Z = (long long)nmod * zs;
    a.ps_rows = (long long)rows_cap * nv;
    a.lu_closed = 1;  // every fold's index comes from the library's kNN,
                      // which lays it out in the closed form
    tune = BfTuneReport();
    if (pinned.threads > 0 && bf_config_ok(pinned, m, variant) &&
        bf_rows(a, pinned) <= rows_cap) {
      cfg = pinned;
    } else {
      long long want = 4000000LL / (nmod > 0 ? nmod : 1);
      const int ntune = (int)std::max(1LL, std::min<long long>(nprob, want));
      const int rc = bf_autotune(a, ntune, st.s, &cfg, &tune);
      if (rc) return set_err(NNGP_ERR_HIP, "autotune launch failed (%d)", rc);
    }
    tuned = true;
    graph_g = -1;
  }
  if (graph_g != g || graph_variant != variant || graph_y0 != want_y0)
    NNGP_TRY(record(g, variant, want_y0));
  NNGP_HIP(hipEventRecord(ev0, st.s));
  NNGP_HIP(hipGraphLaunch(exec, st.s));
  NNGP_HIP(hipEventRecord(ev1, st.s));
  // results: pscore (2 per problem), errd, perrd, sol
  double *hb = hbuf.as<double>();
  double *h_score = hb;
  double *h_err = hb + 2 * (size_t)nprob;
  double *h_perr = h_err + nprob;
  double *h_sol = h_perr + nprob;
  NNGP_HIP(hipMemcpyAsync(h_score, pscore.p, sizeof(double) * 2 * nprob,
                          hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(h_err, errd.p, sizeof(double) * nprob,
                          hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(h_perr, perrd.p, sizeof(double) * nprob,
                          hipMemcpyDeviceToHost, st.s));
  const bool want_sol = det && (det->betaHat || det->ab || det->bBInv);
  if (want_sol)
    NNGP_HIP(hipMemcpyAsync(h_sol, sol.p, sizeof(double) * kSol * nprob,
                            hipMemcpyDeviceToHost, st.s));
  if (want_y0) {
    if (det->y0Hat)
      NNGP_HIP(hipMemcpyAsync(det->y0Hat, y0hat.p,
                              sizeof(double) * (size_t)nprob * nho,
                              hipMemcpyDeviceToHost, st.s));
    if (det->y0HatVar)
      NNGP_HIP(hipMemcpyAsync(det->y0HatVar, y0var.p,
                              sizeof(double) * (size_t)nprob * nho,
                              hipMemcpyDeviceToHost, st.s));
  }
  NNGP_HIP(hipStreamSynchronize(st.s));
  float gms = 0.0f;
  NNGP_HIP(hipEventElapsedTime(&gms, ev0, ev1));
  for (int j = 0; j < nprob; j++) {
    if (h_err[j] != (double)INT_MAX)
      return set_err(NNGP_ERR_NOT_PD,
                     "c++ Rf_error: dpotrf failed\n(fold %d, grid point %d: "
                     "the neighbor covariance of site %.0f is not positive "
                     "definite)", j / g, j % g, h_err[j]);
    if (h_perr[j] != (double)INT_MAX)
      return set_err(NNGP_ERR_NOT_PD,
                     "c++ Rf_error: dpotrf failed\n(fold %d, grid point %d: "
                     "the prediction system of held-out site %.0f is not "
                     "positive definite)", j / g, j % g, h_perr[j]);
  }
  // scores of R/spConjNNGP.R:336-350 (columns rmspe, crps)
  for (int gi = 0; gi < g; gi++) {
    double rm = 0.0, cr = 0.0;
    for (int f = 0; f < k; f++) {
      const int j = f * g + gi;
      cr += h_score[2 * j] / nho;
      rm += std::sqrt(h_score[2 * j + 1] / nho);
    }
    scores[gi] = rm / k;
    scores[(size_t)g + gi] = cr / k;
  }
  if (det) {
    det->graph_ms = gms;
    det->bf_family = cfg.family;
    det->bf_threads = cfg.threads;
    det->bf_mode = cfg.mode;
    det->tune_n = tune.n;
    for (int t = 0; t < NNGP_TUNE_MAX; t++) {
      det->tune_family[t] = t < tune.n ? tune.family[t] : 0;
      det->tune_threads[t] = t < tune.n ? tune.threads[t] : 0;
      det->tune_mode[t] = t < tune.n ? tune.mode[t] : 0;
      det->tune_ms[t] = t < tune.n ? tune.ms[t] : 0.0;
    }
    if (want_sol) {
      for (int j = 0; j < nprob; j++) {
        const double *o = h_sol + (size_t)j * kSol;
        if (!(o[kSolOk] > 0.0))
          return set_err(NNGP_ERR_NOT_PD,
                         "c++ Rf_error: dpotrf failed\n(fold %d, grid point "
                         "%d: X'V^-1X is not positive definite)", j / g,
                         j % g);
        if (det->betaHat)
          for (int r = 0; r < p; r++) det->betaHat[(size_t)j * p + r] = o[r];
        if (det->ab) {
          det->ab[2 * (size_t)j] = o[kSolA];
          det->ab[2 * (size_t)j + 1] = o[kSolB];
        }
        if (det->bBInv)
          for (int r = 0; r < p * p; r++)
            det->bBInv[(size_t)j * p * p + r] = o[kSolV + r];
      }
    }
  }
  return NNGP_OK;
}

}  // namespace nngp

using namespace nngp;

struct nngp_cvgrid {
  CvGrid cv;
};

int nngp_cvgrid_create(nngp_cvgrid **out, const double *y, const double *X,
                       int p, int n, const double *coords, int m, int kFold,
                       const int *hoIndx, int nHo, const double *sigmaSqIG,
                       int covModel, const nngp_opts *opts) {
  clear_err();
  if (!out) return set_err(NNGP_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (covModel < 0 || covModel > 3)
    return set_err(NNGP_ERR_COV_MODEL,
                   "c++ Rf_error: cov.model is not correctly specified");
  if (p < 1 || p > kMaxP)
    return set_err(NNGP_ERR_UNSUPPORTED, "p = %d covariates (1..%d supported)",
                   p, kMaxP);
  if (kFold < 1 || nHo < 1 || n - nHo < 2)
    return set_err(NNGP_ERR_ARG, "need k.fold >= 1 and 1 <= n.ho <= n - 2");
  if (m < 1 || m >= n - nHo)
    return set_err(NNGP_ERR_ARG, "n.neighbors must be in [1, n - n.ho - 1]");
  if (m > 64) return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  if (!y || !X || !coords || !hoIndx || !sigmaSqIG)
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  const int search = opts ? opts->search : 0;
  if (search != 0 && search != 1)
    return set_err(NNGP_ERR_ARG,
                   "search = %d (0: code book, 1: brute) is not defined",
                   search);
  NNGP_TRY(use_device(opts));
  nngp_cvgrid *c = new (std::nothrow) nngp_cvgrid;
  if (!c) return set_err(NNGP_ERR_NOMEM, "host allocation failed");
  const int st = c->cv.create(y, X, p, n, coords, m, kFold, hoIndx, nHo,
                              sigmaSqIG, covModel, search);
  if (st != NNGP_OK) {
    delete c;
    return st;
  }
  if (opts && opts->shape[1] > 0) {
    const BfConfig pin{opts->shape[0], opts->shape[1], opts->shape[2]};
    // a matern grid takes the per-lane shapes at nu = 0.5, 1.5, 2.5; the
    // grid is FP64, so mode 2 (covariate rows staged in LDS) is the only
    // per-lane mode besides 0
    const int variant = covModel == 2 ? kCovMatern15 : covModel;
    if (!bf_config_ok(pin, m, variant) || pin.family == 2 ||  // batched
        (pin.family == 0 && pin.mode != 0 && pin.mode != 2)) {
      delete c;
      return set_err(NNGP_ERR_ARG,
                     "launch shape (%d, %d, %d) does not apply to this grid",
                     pin.family, pin.threads, pin.mode);
    }
    c->cv.pinned = pin;
  }
  *out = c;
  return NNGP_OK;
}

void nngp_cvgrid_destroy(nngp_cvgrid *cv) { delete cv; }

int nngp_cvgrid_run(nngp_cvgrid *cv, const double *thetaAlpha, int g,
                    double *scores, nngp_cvgrid_detail *detail) {
  clear_err();
  if (!cv || !thetaAlpha || !scores || g < 1)
    return set_err(NNGP_ERR_ARG, "NULL argument or g < 1");
  return cv->cv.run(thetaAlpha, g, scores, detail);
}
