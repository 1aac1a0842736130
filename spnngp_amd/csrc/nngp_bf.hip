// nngp_bf.hip — neighbor-system kernels: updateBF / updateConjBF
// (src/rNNGP.cpp:46-101, src/cNNGP.cpp:24-69), the quadratic form Q
// (src/util.cpp:261-280) and the site sums the samplers and the conjugate
// grid need.
//
// Two updateBF kernels, both adding up their sites' terms in the same
// launch (log F_i and e_i e_i^T / F_i, with e_i the residual vector of X
// and y at site i), so one launch gives everything a Metropolis step or a
// conjugate grid point needs:
//  * k_bf_lane<M, COV, BATCH>: one site per lane (64 sites per wavefront),
//    the whole M x M system in registers, fully unrolled (M <= 16,
//    closed-form covariances).  Sites with fewer than M neighbors (the
//    first m sites of the NNGP order) are padded with far-away points:
//    their rows/columns of C are exactly zero off the diagonal, so the real
//    block factors bit-for-bit as if unpadded and the padded coefficients
//    come out exactly 0.  Block size 64, 128 or 256 threads.
//  * k_bf_coop<BATCH, MM>: one site per wavefront, the matrix built in LDS
//    and factored with one row per lane in registers (any m <= MM <= 64,
//    any covariance including the Bessel-K matern),
//    1, 2 or 4 wavefronts per block walking a contiguous run of sites.
// Both compute C = s2 R(N,N) + t2 I, c = s2 R(N,i), the Cholesky factor of C
// (pivots kept as reciprocal square roots), v = L^-1 c, B = L^-T v and
// F = s2 + t2 - v.v, i.e. the reference's potrf -> potri -> symv result
// without forming the inverse.
//
// BATCH = true runs many problems in one launch: gridDim.y problems, problem
// j is fold j / ng at grid point j % ng (the conjugate model's k-fold grid
// search); fold arrays sit at fixed strides and each problem has its own
// covariance parameters, partial rows and failure flag.
//
// Sums are deterministic for a given launch configuration: every thread
// adds a fixed set of sites in a fixed order, each block reduces with a
// fixed butterfly per wavefront and its wavefronts in order into one
// partial row, and k_rows_sum adds the rows of each column in a fixed order.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "nngp_kernels.hpp"
#include "nngp_group.hpp"
#include "nngp_wave.hpp"

namespace nngp {

constexpr int kLaneMaxThreads = 256;
constexpr int kCoopMaxWaves = 4;
constexpr int kCoopMaxM = 64;
constexpr int kCoopBlocks = 2048;  // logical blocks of the wavefront kernel
constexpr int kGrpMaxM = 32;       // the row kernel's largest system
constexpr int kRedThreads = 256;   // threads per block of the sum kernels
constexpr int kRowsThreads = 1024; // threads per column of k_rows_sum
constexpr int kRedBlocks = 1024;   // grid of the sum kernels (at most)

// DPP lane move of a double (two 32-bit moves); lanes outside ROWMASK get 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL,
                                             ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL,
                                             ROWMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}

// The same where lanes outside ROWMASK may receive anything (no value is
// prepared for them): for sums whose result is read in lanes the move
// reaches.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_mov_any(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, ROWMASK,
                                          0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, ROWMASK,
                                          0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes with DPP moves only (no LDS round trips); the
// total is valid in lane 63 (other lanes hold partial or meaningless
// sums).  Fixed order: pairs, quads, half rows, rows, then rows 0+1 and
// 2+3, then all four.
__device__ __forceinline__ double wave_sum_to63(double x) {
  x += dpp_mov_any<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
  x += dpp_mov_any<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
  x += dpp_mov_any<0x141, 0xF>(x);  // row_half_mirror
  x += dpp_mov_any<0x140, 0xF>(x);  // row_mirror
  x += dpp_mov_any<0x142, 0xA>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_mov_any<0x143, 0xC>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

int rows_capacity(int n) {
  const int lane_rows = (n + 63) / 64;  // smallest lane block is 64 threads
  return lane_rows > kCoopBlocks ? lane_rows : kCoopBlocks;
}

// ---------------------------------------------------------------------------
// problem view
// ---------------------------------------------------------------------------
struct BfProblem {
  const double2 *xy;
  const int *nnIndx;
  const int *lu;
  const double *Z;
  double *rows;
  int *err;
};

template <bool BATCH>
__device__ __forceinline__ BfProblem bf_problem(const BfArgs &A) {
  BfProblem P{A.xy, A.nnIndx, A.lu, A.Z, A.rows, A.err};
  if constexpr (BATCH) {
    const int j = (int)blockIdx.y;
    const int f = j / A.ng;
    P.xy += (size_t)f * A.fs_xy;
    P.nnIndx += (size_t)f * A.fs_nn;
    P.lu += (size_t)f * A.fs_lu;
    if (P.Z) P.Z += (size_t)f * A.fs_Z;
    if (P.rows) P.rows += (size_t)j * A.ps_rows;
    P.err += j;
  }
  return P;
}

// ---------------------------------------------------------------------------
// one site per lane
// ---------------------------------------------------------------------------

// The block's partial row: column 0 sum log F_i, column 1 zero (quadratic
// forms are formed from the cross-products), then the lower triangle of
// e e^T / F row by row.  Called by every thread of the block.
__device__ inline void block_site_terms(int nz, double *rows, int rows_cap,
                                        bool valid, double F,
                                        const double (&e)[kMaxNZ], int row) {
  __shared__ double red[kLaneMaxThreads / 64][kStatsMax];
  const int nw = (int)blockDim.x >> 6;
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  const double rf = valid ? 1.0 / F : 0.0;
  double x = wave_sum_to63(valid ? log(F) : 0.0);
  if (t == 63) red[w][0] = x;
  int g = 2;
#pragma unroll
  for (int p = 0; p < kMaxNZ; p++) {
    if (p < nz) {
#pragma unroll
      for (int q = 0; q <= p; q++) {
        x = wave_sum_to63(e[p] * e[q] * rf);
        if (t == 63) red[w][g + q] = x;
      }
    }
    g += p + 1;
  }
  __syncthreads();
  const int nv = 2 + nz * (nz + 1) / 2;
  const int k = (int)threadIdx.x;
  if (k < nv) {
    double s = 0.0;
    if (k != 1)
      for (int ww = 0; ww < nw; ww++) s += red[ww][k];
    rows[(size_t)k * rows_cap + row] = s;
  }
}

// Number of per-site terms of a partial row (column 1, always 0, aside):
// log F, then the nz(nz+1)/2 cross-products.
__host__ __device__ constexpr int site_terms(int nz) {
  return 1 + nz * (nz + 1) / 2;
}

// block_site_terms through LDS.  Each lane writes its terms to its wave's
// area `w` (term v of lane t at w[v * 64 + t], site_terms(nz) x 64 doubles),
// each quad of lanes then sums one term over 16 lanes apiece in lane order
// and combines its four partials (pairs, then the two pairs), and with more
// than one wave per block the wave sums meet in LDS and are added in wave
// order.  The order is fixed, so a launch shape's sums are reproducible.
// `w` may be memory the wave has finished reading (the staged covariate
// rows); the barrier before the writes also keeps those reads first.
__device__ inline void block_site_terms_lds(int nz, double *rows,
                                            int rows_cap, bool valid,
                                            double F,
                                            const double (&e)[kMaxNZ],
                                            int row, double *w) {
  __shared__ double wsum[kLaneMaxThreads / 64][kStatsMax];
  const int nw = (int)blockDim.x >> 6;
  const int wv = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  const int nt = site_terms(nz);
  const double rf = valid ? 1.0 / F : 0.0;
  __syncthreads();
  w[t] = valid ? log(F) : 0.0;
  int g = 1;
#pragma unroll
  for (int p = 0; p < kMaxNZ; p++) {
    if (p < nz) {
#pragma unroll
      for (int q = 0; q <= p; q++) w[(g + q) * 64 + t] = e[p] * e[q] * rf;
    }
    g += p + 1;
  }
  __syncthreads();
  for (int v0 = 0; v0 < nt; v0 += 16) {
    const int v = v0 + (t >> 2), qd = t & 3;
    double s = 0.0;
    if (v < nt) {
      const double2 *src = reinterpret_cast<const double2 *>(w + v * 64 +
                                                             qd * 16);
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const double2 x = src[u];
        s += x.x;
        s += x.y;
      }
    }
    s += dpp_mov_any<0xB1, 0xF>(s);  // quad_perm [1,0,3,2]
    s += dpp_mov_any<0x4E, 0xF>(s);  // quad_perm [2,3,0,1]
    if (qd == 0 && v < nt) {
      if (nw == 1)
        rows[(size_t)(v == 0 ? 0 : v + 1) * rows_cap + row] = s;
      else
        wsum[wv][v] = s;
    }
  }
  if (nw == 1) {
    if (t == 0) rows[(size_t)rows_cap + row] = 0.0;
    return;
  }
  __syncthreads();
  const int k = (int)threadIdx.x;
  if (k <= nt) {  // column k of the row: term k - 1 beyond column 1
    double s = 0.0;
    if (k != 1) {
      const int v = k == 0 ? 0 : k - 1;
      for (int ww = 0; ww < nw; ww++) s += wsum[ww][v];
    }
    rows[(size_t)k * rows_cap + row] = s;
  }
}

// Direct-to-LDS load of 16 bytes a lane: lane t's bytes land at dst + t
// (dst the same for the whole wavefront).  The load counts in vmcnt like any
// other; the data are in LDS once it has been waited for.
__device__ __forceinline__ void lds_load16(const double *src, double2 *dst) {
  __builtin_amdgcn_global_load_lds(
      (const __attribute__((address_space(1))) void *)src,
      (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

// One site per lane: B_i, F_i and the residual vector e of site i from its
// neighbor count, offset and index slots (slots past the count are not its
// neighbors; they are replaced by site 0 here, before any use).
// ZH > 0 stages the rows of Z (zs = 2 ZH doubles) of the M neighbor slots
// and of the site itself in LDS with the coordinate gathers, instead of
// gathering them into registers after the solve, where the wait for them
// would stall the wavefront: chunk h of slot k (slot M = the site) at
// stg[(k * ZH + h) * 64 + lane].  The residuals add the same terms in the
// same order either way.
template <int M, int COV, bool BATCH, int ZH = 0>
__device__ __forceinline__ void lane_site(const BfArgs &A, const BfProblem &P,
                                          const CovArgs &c, int i, int cnt,
                                          int off, int (&nb)[M], double &Fi,
                                          double (&e)[kMaxNZ],
                                          double2 *stg = nullptr) {
  if (cnt == 0) {  // site 0: F = s2 + t2, residual = the data
    Fi = c.diag;
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++)
      if (j < A.nz) e[j] = P.Z[(size_t)i * A.zs + j];
    return;
  }
  if (cnt < M) {
#pragma unroll
    for (int k = 0; k < M; k++)
      if (k >= cnt) nb[k] = 0;
  }
  // The staged rows are requested before the coordinates: the hardware's
  // load counter retires in order, so the one wait for the coordinates
  // covers them too, and nothing waits on memory again until the residuals.
  if constexpr (ZH > 0) {
#pragma unroll
    for (int k = 0; k <= M; k++) {
      const double *zr = P.Z + (size_t)(k < M ? nb[k] : i) * (2 * ZH);
#pragma unroll
      for (int h = 0; h < ZH; h++)
        lds_load16(zr + 2 * h, stg + (k * ZH + h) * 64);
    }
  }
  // the site's own coordinates first, so that they arrive with the
  // neighbors' (one wait for both)
  const double2 s = P.xy[i];
  double px[M], py[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const double2 q = P.xy[nb[k]];
    px[k] = q.x;
    py[k] = q.y;
  }
  if (cnt < M) {
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k >= cnt) {  // far away and far apart: zero correlation
        px[k] = 1e100 * (k + 1);
        py[k] = -1e100 * (k + 1);
      }
    }
  }
  // exponential, gaussian and the closed-form materns: neighbor offsets
  // from the site, in units where the exponential factor is 2^-d (2^-d^2),
  // see s2_exp2.  Offsetting first keeps large coordinate values (projected
  // systems) out of the rounding of the scaled coordinates.
  constexpr bool kBase2 = lane_base2<COV>();
  if constexpr (kBase2) {
#pragma unroll
    for (int k = 0; k < M; k++) {
      px[k] = (px[k] - s.x) * c.tsc;
      py[k] = (py[k] - s.y) * c.tsc;
    }
  }

  constexpr int T = M * (M + 1) / 2;
  double a[T];  // packed lower triangle
  double v[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    v[k] = lane_corr<COV>(kBase2 ? lane_d2<COV>(px[k], py[k])
                                 : sqdist(s.x, s.y, px[k], py[k]),
                          c);
#pragma unroll
    for (int l = 0; l < k; l++)
      a[tri(k, l)] = lane_corr<COV>(
          lane_d2<COV>(px[k] - px[l], py[k] - py[l]), c);
    a[tri(k, k)] = c.diag;
  }
  double vv;
  const bool ok = chol_solve<M>(a, v, vv);  // v <- B_i = C^-1 c
  if (!BATCH && A.B) {
#pragma unroll
    for (int k = 0; k < M; k++)
      if (k < cnt) A.B[off + k] = v[k];
  }
  Fi = c.diag - vv;
  // residual vectors; padded coefficients are exactly 0
  if constexpr (ZH > 0) {
    const int t = (int)threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < ZH; h++) {
      const double2 zi = stg[(M * ZH + h) * 64 + t];
      double e0 = zi.x, e1 = zi.y;
#pragma unroll
      for (int k = 0; k < M; k++) {
        const double2 z = stg[(k * ZH + h) * 64 + t];
        e0 = fma(-v[k], z.x, e0);
        e1 = fma(-v[k], z.y, e1);
      }
      e[2 * h] = e0;
      e[2 * h + 1] = e1;
    }
    if (!ok) atomicMin(P.err, i);
    return;
  }
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) {
    if (j < A.nz) {
      double ej = P.Z[(size_t)i * A.zs + j];
#pragma unroll
      for (int k = 0; k < M; k++)
        ej = fma(-v[k], P.Z[(size_t)nb[k] * A.zs + j], ej);
      e[j] = ej;
    }
  }
  if (!ok) atomicMin(P.err, i);
}

template <int M, int COV, bool BATCH>
__global__ __launch_bounds__(kLaneMaxThreads) void k_bf_lane(BfArgs A) {
  extern __shared__ double lane_red[];  // site_terms(nz) x 64 per wave
  const int nt = (int)blockDim.x;
  const int b = xcd_block((A.n + nt - 1) / nt);
  if (b < 0) return;  // whole padding block
  const BfProblem P = bf_problem<BATCH>(A);
  const CovArgs c = BATCH ? A.covs[(int)blockIdx.y % A.ng] : A.c;
  const int i = b * nt + (int)threadIdx.x;
  const bool valid = i < A.n;
  double Fi = 1.0;
  double e[kMaxNZ];
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;

  if (valid) {
    int cnt, off;
    if (A.lu_closed) {
      nn_closed(i, A.m, off, cnt);
    } else {
      cnt = P.lu[A.n + i];
      off = P.lu[i];
    }
    // The index array carries 64 padding entries, so off + k is in bounds.
    int nb[M];
#pragma unroll
    for (int k = 0; k < M; k++) nb[k] = P.nnIndx[off + k];
    lane_site<M, COV, BATCH>(A, P, c, i, cnt, off, nb, Fi, e);
    if (!BATCH && A.F) A.F[i] = Fi;
  }
  if (P.rows) {
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    block_site_terms_lds(A.nz, P.rows, A.rows_cap, valid, Fi, e, b,
                         lane_red + (size_t)w * site_terms(A.nz) * 64);
  }
}

// k_bf_lane with the covariate rows staged in LDS: the block's dynamic LDS
// holds, per wavefront, (M + 1) x ZH chunks of 64 x 16 bytes.  The launcher
// uses it only when Z is set, zs = 2 ZH and every problem's Z is 16-byte
// aligned.
template <int M, int COV, int ZH, bool BATCH>
__global__ __launch_bounds__(kLaneMaxThreads) void k_bf_lane_lds(BfArgs A) {
  extern __shared__ double2 zstage[];
  const int nt = (int)blockDim.x;
  const int b = xcd_block((A.n + nt - 1) / nt);
  if (b < 0) return;  // whole padding block
  const BfProblem P = bf_problem<BATCH>(A);
  const CovArgs c = BATCH ? A.covs[(int)blockIdx.y % A.ng] : A.c;
  const int i = b * nt + (int)threadIdx.x;
  const bool valid = i < A.n;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  double2 *stg = zstage + (size_t)w * (M + 1) * ZH * 64;
  double Fi = 1.0;
  double e[kMaxNZ];
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;

  if (valid) {
    int cnt, off;
    if (A.lu_closed) {
      nn_closed(i, A.m, off, cnt);
    } else {
      cnt = P.lu[A.n + i];
      off = P.lu[i];
    }
    int nb[M];
#pragma unroll
    for (int k = 0; k < M; k++) nb[k] = P.nnIndx[off + k];
    lane_site<M, COV, BATCH, ZH>(A, P, c, i, cnt, off, nb, Fi, e, stg);
    if (!BATCH && A.F) A.F[i] = Fi;
  }
  // the sums reuse the wave's staging area (the residuals have read it);
  // site_terms(nz) <= 11 doubles a lane fit in (M + 1) x ZH x 16 bytes
  if (P.rows)
    block_site_terms_lds(A.nz, P.rows, A.rows_cap, valid, Fi, e, b,
                         reinterpret_cast<double *>(stg));
}

// ---------------------------------------------------------------------------
// mixed precision, one site per lane
// ---------------------------------------------------------------------------
// The covariance is built in FP32 from FP32 neighbor offsets (FP64
// differences to the site, rounded), factored and solved in FP32 for B0.
// One refinement step forms the residual r = c - C B0 of that FP32 system
// in FP64 arithmetic (each FP32 x FP64 product is exact before the sum),
// solves C d = r with the FP32 factor and sets B = B0 + d, which removes
// the factorization's error.  F = s2 + t2 - c.B and every sum stay in FP64.
// What remains is the FP32 rounding of the covariance entries themselves.
// COPY = true keeps C in LDS (lane-interleaved) for the residual; false
// evaluates the entries a second time (bit-identically) to save LDS.
template <int COV>
__device__ __forceinline__ float pair32(float dxk, float dyk, float dxl,
                                        float dyl, const CovArgs32 &c) {
  const float ex = dxk - dxl, ey = dyk - dyl;
  return s2_corr32<COV>(fmaf(ex, ex, ey * ey), c);
}

template <int M, int COV, bool COPY>
__global__ __launch_bounds__(kLaneMaxThreads, 2) void k_bf_mixed(BfArgs A) {
  // COPY: covariance entry t of lane L at cl[t*nt + L]; otherwise the
  // neighbor offsets dx_k at cl[k*nt + L], dy_k at cl[(M+k)*nt + L]
  extern __shared__ float cl[];
  const int nt = (int)blockDim.x;
  const int b = xcd_block((A.n + nt - 1) / nt);
  if (b < 0) return;
  const CovArgs &c = A.c;
  const CovArgs32 c32 = cov32(c);
  const int L = (int)threadIdx.x;
  const int i = b * nt + L;
  const bool valid = i < A.n;
  double Fi = 1.0;
  double e[kMaxNZ];  // set on every path just before use (no value is live
                     // through the factorization)

  if (valid) {
    int cnt, off;
    if (A.lu_closed) {
      nn_closed(i, A.m, off, cnt);
    } else {
      cnt = A.lu[A.n + i];
      off = A.lu[i];
    }
    if (cnt == 0) {
      Fi = c.diag;
#pragma unroll
      for (int j = 0; j < kMaxNZ; j++)
        e[j] = j < A.nz ? A.Z[(size_t)i * A.zs + j] : 0.0;
    } else {
      const double2 s = A.xy[i];
      // as in k_bf_lane: padding entries of the index read as site 0, then
      // replaced by far-away finite offsets (every product with one is +inf
      // or 0 in FP32, so its correlations are exactly 0)
      int nb[M];
#pragma unroll
      for (int k = 0; k < M; k++) nb[k] = A.nnIndx[off + k];
      if (cnt < M) {
#pragma unroll
        for (int k = 0; k < M; k++)
          if (k >= cnt) nb[k] = 0;
      }
      float dx[M], dy[M];
#pragma unroll
      for (int k = 0; k < M; k++) {
        const double2 q = A.xy[nb[k]];
        dx[k] = (float)(q.x - s.x);
        dy[k] = (float)(q.y - s.y);
      }
      if (cnt < M) {
#pragma unroll
        for (int k = 0; k < M; k++) {
          if (k >= cnt) {
            dx[k] = 1e30f * (k + 1);
            dy[k] = -1e30f * (k + 1);
          }
        }
      }
      if (!COPY) {  // read back by the residual; registers stay free
#pragma unroll
        for (int k = 0; k < M; k++) {
          cl[k * nt + L] = dx[k];
          cl[(M + k) * nt + L] = dy[k];
        }
      }
      constexpr int T = M * (M + 1) / 2;
      float a[T], v[M];
#pragma unroll
      for (int k = 0; k < M; k++) {
        if (COPY) {
          v[k] = s2_corr32<COV>(fmaf(dx[k], dx[k], dy[k] * dy[k]), c32);
          cl[(T + k) * nt + L] = v[k];
        }
#pragma unroll
        for (int l = 0; l < k; l++) {
          a[tri(k, l)] = pair32<COV>(dx[k], dy[k], dx[l], dy[l], c32);
          if (COPY) cl[tri(k, l) * nt + L] = a[tri(k, l)];
        }
        a[tri(k, k)] = c32.diag;
      }
      const bool ok = chol_factor32<M>(a);
      if (!COPY) {  // c after the factorization, from the offsets in LDS
#pragma unroll
        for (int k = 0; k < M; k++) {
          const float xk = cl[k * nt + L], yk = cl[(M + k) * nt + L];
          v[k] = s2_corr32<COV>(fmaf(xk, xk, yk * yk), c32);
        }
      }
      chol_apply32<M>(a, v);  // v <- B0
      // residual r = c - C B0 in FP64; B0 stays FP32 (converted where
      // used), so only r is an FP64 array while the factor is live
      double r[M];
      const double dg = (double)c32.diag;
#pragma unroll
      for (int k = 0; k < M; k++) {
        float ck;
        if (COPY) {
          ck = cl[(T + k) * nt + L];
        } else {
          const float xk = cl[k * nt + L], yk = cl[(M + k) * nt + L];
          ck = s2_corr32<COV>(fmaf(xk, xk, yk * yk), c32);
        }
        r[k] = fma(-dg, (double)v[k], (double)ck);
      }
#pragma unroll
      for (int k = 1; k < M; k++) {
        float xk = 0.0f, yk = 0.0f;
        if (!COPY) {
          xk = cl[k * nt + L];
          yk = cl[(M + k) * nt + L];
        }
#pragma unroll
        for (int l = 0; l < k; l++) {
          const double ckl =
              COPY ? (double)cl[tri(k, l) * nt + L]
                   : (double)pair32<COV>(xk, yk, cl[l * nt + L],
                                         cl[(M + l) * nt + L], c32);
          r[k] = fma(-ckl, (double)v[l], r[k]);
          r[l] = fma(-ckl, (double)v[k], r[l]);
        }
      }
      float d[M];
#pragma unroll
      for (int k = 0; k < M; k++) d[k] = (float)r[k];
      chol_apply32<M>(a, d);  // d <- C^-1 r
      double bd[M];
      double cb = 0.0;
#pragma unroll
      for (int k = 0; k < M; k++) {
        bd[k] = (double)v[k] + (double)d[k];
        float ck;
        if (COPY) {
          ck = cl[(T + k) * nt + L];
        } else {
          const float xk = cl[k * nt + L], yk = cl[(M + k) * nt + L];
          ck = s2_corr32<COV>(fmaf(xk, xk, yk * yk), c32);
        }
        cb = fma((double)ck, bd[k], cb);
      }
      if (A.B) {
#pragma unroll
        for (int k = 0; k < M; k++)
          if (k < cnt) A.B[off + k] = bd[k];
      }
      Fi = c.diag - cb;
      // the neighbor slots again (cache hits) rather than held in
      // registers through the factorization; padded coefficients are 0
#pragma unroll
      for (int k = 0; k < M; k++) nb[k] = k < cnt ? A.nnIndx[off + k] : 0;
#pragma unroll
      for (int j = 0; j < kMaxNZ; j++) {
        e[j] = 0.0;
        if (j < A.nz) {
          double ej = A.Z[(size_t)i * A.zs + j];
#pragma unroll
          for (int k = 0; k < M; k++)
            ej = fma(-bd[k], A.Z[(size_t)nb[k] * A.zs + j], ej);
          e[j] = ej;
        }
      }
      if (!ok || !(Fi > 0.0)) atomicMin(A.err, i);
    }
    if (A.F) A.F[i] = Fi;
  } else {
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;
  }
  if (A.rows) block_site_terms(A.nz, A.rows, A.rows_cap, valid, Fi, e, b);
}

// k_bf_mixed's scheme at one wave per SIMD, where the FP32 factor and a
// register copy of C both fit: no entry is evaluated twice and C does not
// go through LDS.  The arithmetic is k_bf_mixed's step for step, so B and F
// equal its results bit for bit.  As in lane_site, ZH > 0 stages the
// covariate rows of the M neighbor slots and of the site in LDS with the
// coordinate gathers (chunk h of slot k at stg[(k * ZH + h) * 64 + lane]).
template <int M, int COV, int ZH>
__device__ __forceinline__ void mixed_reg_site(const BfArgs &A, int i,
                                               int cnt, int off,
                                               int (&nb)[M], double &Fi,
                                               double (&e)[kMaxNZ],
                                               double2 *stg) {
  const CovArgs &c = A.c;
  const CovArgs32 c32 = cov32(c);
  if (cnt == 0) {  // site 0: F = s2 + t2, residual = the data
    Fi = c.diag;
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++)
      if (j < A.nz) e[j] = A.Z[(size_t)i * A.zs + j];
    return;
  }
  if (cnt < M) {
#pragma unroll
    for (int k = 0; k < M; k++)
      if (k >= cnt) nb[k] = 0;
  }
  if constexpr (ZH > 0) {
#pragma unroll
    for (int k = 0; k <= M; k++) {
      const double *zr = A.Z + (size_t)(k < M ? nb[k] : i) * (2 * ZH);
#pragma unroll
      for (int h = 0; h < ZH; h++)
        lds_load16(zr + 2 * h, stg + (k * ZH + h) * 64);
    }
  }
  const double2 s = A.xy[i];
  float dx[M], dy[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const double2 q = A.xy[nb[k]];
    dx[k] = (float)(q.x - s.x);
    dy[k] = (float)(q.y - s.y);
  }
  if (cnt < M) {  // far away and far apart: correlations exactly 0
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k >= cnt) {
        dx[k] = 1e30f * (k + 1);
        dy[k] = -1e30f * (k + 1);
      }
    }
  }
  constexpr int T = M * (M + 1) / 2;
  float a[T], ac[T], v[M], cv[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    v[k] = s2_corr32<COV>(fmaf(dx[k], dx[k], dy[k] * dy[k]), c32);
    cv[k] = v[k];
#pragma unroll
    for (int l = 0; l < k; l++) {
      a[tri(k, l)] = pair32<COV>(dx[k], dy[k], dx[l], dy[l], c32);
      ac[tri(k, l)] = a[tri(k, l)];
    }
    a[tri(k, k)] = c32.diag;
  }
  const bool ok = chol_factor32<M>(a);
  chol_apply32<M>(a, v);  // v <- B0
  // residual r = c - C B0 in FP64
  double r[M];
  const double dg = (double)c32.diag;
#pragma unroll
  for (int k = 0; k < M; k++) r[k] = fma(-dg, (double)v[k], (double)cv[k]);
#pragma unroll
  for (int k = 1; k < M; k++) {
#pragma unroll
    for (int l = 0; l < k; l++) {
      const double ckl = (double)ac[tri(k, l)];
      r[k] = fma(-ckl, (double)v[l], r[k]);
      r[l] = fma(-ckl, (double)v[k], r[l]);
    }
  }
  float d[M];
#pragma unroll
  for (int k = 0; k < M; k++) d[k] = (float)r[k];
  chol_apply32<M>(a, d);  // d <- C^-1 r
  double bd[M];
  double cb = 0.0;
#pragma unroll
  for (int k = 0; k < M; k++) {
    bd[k] = (double)v[k] + (double)d[k];
    cb = fma((double)cv[k], bd[k], cb);
  }
  if (A.B) {
#pragma unroll
    for (int k = 0; k < M; k++)
      if (k < cnt) A.B[off + k] = bd[k];
  }
  Fi = c.diag - cb;
  // residual vectors; padded coefficients are 0
  if constexpr (ZH > 0) {
    const int t = (int)threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < ZH; h++) {
      const double2 zi = stg[(M * ZH + h) * 64 + t];
      double e0 = zi.x, e1 = zi.y;
#pragma unroll
      for (int k = 0; k < M; k++) {
        const double2 z = stg[(k * ZH + h) * 64 + t];
        e0 = fma(-bd[k], z.x, e0);
        e1 = fma(-bd[k], z.y, e1);
      }
      e[2 * h] = e0;
      e[2 * h + 1] = e1;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++) {
      if (j < A.nz) {
        double ej = A.Z[(size_t)i * A.zs + j];
#pragma unroll
        for (int k = 0; k < M; k++)
          ej = fma(-bd[k], A.Z[(size_t)nb[k] * A.zs + j], ej);
        e[j] = ej;
      }
    }
  }
  if (!ok || !(Fi > 0.0)) atomicMin(A.err, i);
}

// Dynamic LDS of k_bf_mixed_reg a wavefront, in double2: the staged rows,
// which the block sums then reuse, or the sums' area alone.
__host__ __device__ inline int mixed_reg_wave_chunks(int M, int ZH, int nz) {
  const int sums = (site_terms(nz) * 64 + 1) / 2;
  const int stage = (M + 1) * ZH * 64;
  return stage > sums ? stage : sums;
}

template <int M, int COV, int ZH>
__global__ __launch_bounds__(kLaneMaxThreads) void k_bf_mixed_reg(BfArgs A) {
  extern __shared__ double2 mstage[];
  const int nt = (int)blockDim.x;
  const int b = xcd_block((A.n + nt - 1) / nt);
  if (b < 0) return;  // whole padding block
  const int i = b * nt + (int)threadIdx.x;
  const bool valid = i < A.n;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  double2 *stg = mstage + (size_t)w * mixed_reg_wave_chunks(M, ZH, A.nz);
  double Fi = 1.0;
  double e[kMaxNZ];
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;
  if (valid) {
    int cnt, off;
    if (A.lu_closed) {
      nn_closed(i, A.m, off, cnt);
    } else {
      cnt = A.lu[A.n + i];
      off = A.lu[i];
    }
    int nb[M];
#pragma unroll
    for (int k = 0; k < M; k++) nb[k] = A.nnIndx[off + k];
    mixed_reg_site<M, COV, ZH>(A, i, cnt, off, nb, Fi, e, stg);
    if (A.F) A.F[i] = Fi;
  }
  if (A.rows)
    block_site_terms_lds(A.nz, A.rows, A.rows_cap, valid, Fi, e, b,
                         reinterpret_cast<double *>(stg));
}

// ---------------------------------------------------------------------------
// one site per wavefront (generic m, generic covariance)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tri_row(int p) {
  // largest k with k(k+1)/2 <= p
  int k = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
  while ((k + 1) * (k + 2) / 2 <= p) k++;
  while (k * (k + 1) / 2 > p) k--;
  return k;
}

// Per-wave LDS of the wavefront kernels, in doubles.
__host__ __device__ inline int coop_wave_doubles(int m) {
  return m * (m + 1) + 3 * m;
}

// Solve site i's system in the wave's LDS area Am; every lane returns F and
// the residual vector e (identical in all lanes); lane t < cnt returns its
// coefficient B_t in bcoef.  False when a pivot is not positive.
template <int MM>
__device__ inline bool coop_site(const BfArgs &A, const BfProblem &P,
                                 const CovArgs &c, const BesselTab &bt,
                                 const MaternTab &mt, double *Am, int i,
                                 double &F,
                                 double (&e)[kMaxNZ], double &bcoef,
                                 int &cnt_out, int &off_out) {
  const int t = lane_opaque((int)threadIdx.x & 63);
  const int ld = A.m + 1;
  double *cv = Am + A.m * ld;
  double *px = cv + A.m;
  double *py = px + A.m;
  // wave-uniform (one site per wavefront): scalar registers, so the
  // guards on cnt below are scalar branches, not exec-mask changes
  const int cnt = __builtin_amdgcn_readfirstlane(P.lu[A.n + i]);
  const int off = __builtin_amdgcn_readfirstlane(P.lu[i]);
  cnt_out = cnt;
  off_out = off;
  bcoef = 0.0;
  if (cnt == 0) {
    F = c.diag;
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++)
      e[j] = j < A.nz ? P.Z[(size_t)i * A.zs + j] : 0.0;
    return true;
  }
  const double2 s = P.xy[i];
  int myid = 0;
  if (t < cnt) {
    myid = P.nnIndx[off + t];
    const double2 q = P.xy[myid];
    px[t] = q.x;
    py[t] = q.y;
  }
  wave_sync();
  // augmented lower triangle: rows 0..cnt-1 are C, row cnt is c
  const int np = (cnt + 1) * (cnt + 2) / 2 - 1;
  for (int p = t; p < np; p += 64) {
    const int k = tri_row(p);
    const int l = p - k * (k + 1) / 2;
    if (k < cnt && k == l) {
      Am[k * ld + k] = c.diag;
    } else {
      const double ax = k < cnt ? px[k] : s.x;
      const double ay = k < cnt ? py[k] : s.y;
      const double val = s2_corr_rt(sqdist(ax, ay, px[l], py[l]), c, bt, mt);
      if (k < cnt)
        Am[k * ld + l] = val;
      else
        cv[l] = val;
    }
  }
  wave_sync();
  double x, vv;
  if (!wave_chol_rows<MM>(Am, ld, cnt, t, t < cnt ? cv[t] : 0.0, x, vv)) {
    F = 1.0;  // keeps the sums finite; the pass reports the failure
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;
    wave_sync();
    return false;
  }
  bcoef = t < cnt ? x : 0.0;
  F = c.diag - vv;
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) {
    e[j] = 0.0;
    if (j < A.nz) {
      const double part =
          wave_sum(t < cnt ? x * P.Z[(size_t)myid * A.zs + j] : 0.0);
      e[j] = P.Z[(size_t)i * A.zs + j] - part;
    }
  }
  wave_sync();  // this LDS area is reused by the wave's next site
  return true;
}

// Per-wave LDS of the mixed wavefront kernel, in floats: C's lower triangle
// and then its factor (m rows of m + 1), the vector c and the neighbor
// offsets.
__host__ __device__ inline int coop_mixed_wave_floats(int m) {
  return m * (m + 1) + 3 * m;
}

// coop_site in mixed precision (the per-lane k_bf_mixed's scheme, one site
// per wavefront): neighbor offsets from site i and the covariance (any
// covModel, the matern by xnu_bessel_k32) in FP32, the FP32 factor one row
// per lane in registers, B0 = C32^-1 c32 in FP32, the residual
// r = c32 - C32 B0 of that FP32 system in FP64 (products of FP32 values are
// exact in FP64; lane t keeps its whole row of C32 in registers for it),
// one FP32 correction solve B = B0 + C32^-1 r, and F = s2 + t2 - c32.B and
// the residuals e in FP64.  False when a pivot or F is not positive.
template <int MM>
__device__ inline bool coop_site_mixed(const BfArgs &A, const BfProblem &P,
                                       const CovArgs &c,
                                       const CovArgs32 &c32,
                                       const BesselTab32 &bt, float *W, int i,
                                       double &F, double (&e)[kMaxNZ],
                                       double &bcoef, int &cnt_out,
                                       int &off_out) {
  const int t = lane_opaque((int)threadIdx.x & 63);
  const int ld = A.m + 1;
  float *Lm = W;
  float *cv = W + A.m * ld;
  float *px = cv + A.m;
  float *py = px + A.m;
  // wave-uniform (one site per wavefront): scalar registers, so the
  // guards on cnt below are scalar branches, not exec-mask changes
  const int cnt = __builtin_amdgcn_readfirstlane(P.lu[A.n + i]);
  const int off = __builtin_amdgcn_readfirstlane(P.lu[i]);
  cnt_out = cnt;
  off_out = off;
  bcoef = 0.0;
  if (cnt == 0) {
    F = c.diag;
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++)
      e[j] = j < A.nz ? P.Z[(size_t)i * A.zs + j] : 0.0;
    return true;
  }
  const double2 s = P.xy[i];
  int myid = 0;
  if (t < cnt) {
    myid = P.nnIndx[off + t];
    const double2 q = P.xy[myid];
    px[t] = (float)(q.x - s.x);
    py[t] = (float)(q.y - s.y);
  }
  wave_sync();
  // augmented lower triangle: rows 0..cnt-1 are C, row cnt is c (site i
  // sits at the origin of the offsets)
  const int np = (cnt + 1) * (cnt + 2) / 2 - 1;
  for (int p = t; p < np; p += 64) {
    const int k = tri_row(p);
    const int l = p - k * (k + 1) / 2;
    if (k < cnt && k == l) {
      Lm[k * ld + k] = c32.diag;
    } else {
      const float ex = (k < cnt ? px[k] : 0.0f) - px[l];
      const float ey = (k < cnt ? py[k] : 0.0f) - py[l];
      const float val = s2_corr32_rt(fmaf(ex, ex, ey * ey), c32, bt);
      if (k < cnt)
        Lm[k * ld + l] = val;
      else
        cv[l] = val;
    }
  }
  wave_sync();
  float crow[MM], row[MM];
#pragma unroll
  for (int l = 0; l < MM; l++) {
    crow[l] = (t < cnt && l < cnt)
                  ? (l <= t ? Lm[t * ld + l] : Lm[l * ld + t])
                  : 0.0f;
    row[l] = l <= t ? crow[l] : 0.0f;
  }
  const float ct = t < cnt ? cv[t] : 0.0f;
  wave_sync();  // C is in registers now; the factor takes its place
  bool ok = wave_factor32<MM>(row, cnt, t);
  double x = 0.0;
  if (ok) {
#pragma unroll
    for (int l = 0; l < MM; l++)
      if (t < cnt && l <= t) Lm[t * ld + l] = row[l];
    wave_sync();
    const float x0 = wave_solve32<MM>(row, Lm, ld, cnt, t, ct);
    double r = (double)ct;
#pragma unroll
    for (int l = 0; l < MM; l++)
      if (l < cnt) r = fma(-(double)crow[l], (double)readlane_f(x0, l), r);
    const float d =
        wave_solve32<MM>(row, Lm, ld, cnt, t, t < cnt ? (float)r : 0.0f);
    x = t < cnt ? (double)x0 + (double)d : 0.0;
    F = c.diag - wave_sum((double)ct * x);
    ok = F > 0.0;
  }
  if (!ok) {
    F = 1.0;  // keeps the sums finite; the pass reports the failure
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;
    wave_sync();
    return false;
  }
  bcoef = x;
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) {
    e[j] = 0.0;
    if (j < A.nz) {
      const double part =
          wave_sum(t < cnt ? x * P.Z[(size_t)myid * A.zs + j] : 0.0);
      e[j] = P.Z[(size_t)i * A.zs + j] - part;
    }
  }
  wave_sync();  // this LDS area is reused by the wave's next site
  return true;
}

// The waves of the logical blocks of one XCD take that XCD's run of sites
// round robin (xcd_site_run).  A.coop_blocks logical blocks per problem.
// MIXED: the sites are solved by coop_site_mixed (single problem only).
// The block's matern tables: FP64 passes the Bessel-K loop divisors and the
// pass's correlation table, mixed passes the FP32 divisors.
template <bool MIXED>
struct CoopTabs {
  BesselTab bt;
  MaternTab mt;
};
template <>
struct CoopTabs<true> {
  BesselTab32 bt32;
};

template <bool BATCH, int MM, bool MIXED>
__global__ __launch_bounds__(64 * kCoopMaxWaves) void k_bf_coop(BfArgs A) {
  extern __shared__ double smem[];
  __shared__ double red[kCoopMaxWaves][kStatsMax];
  __shared__ CoopTabs<MIXED> tabs;
  const int nw = (int)blockDim.x >> 6;
  const int w = (int)threadIdx.x >> 6;
  const int t = (int)threadIdx.x & 63;
  const int b = xcd_block(A.coop_blocks);
  if (b < 0) return;
  const BfProblem P = bf_problem<BATCH>(A);
  const CovArgs c = BATCH ? A.covs[(int)blockIdx.y % A.ng] : A.c;
  if (c.cov == 2) {  // the matern's loop divisors for this block's mu
    if constexpr (MIXED) {
      bessel_tab_fill32(tabs.bt32, c);
      __syncthreads();
    } else {
      bessel_tab_fill(tabs.bt, c);
      __syncthreads();
      // and the pass's correlation table, built in the waves' system areas
      // (at least kMatScratch doubles, see the launcher) before their use
      matern_tab_fill(tabs.mt, c, tabs.bt, smem);
    }
  }
  double *Am = smem + (size_t)w * coop_wave_doubles(A.m);
  float *Wf = reinterpret_cast<float *>(smem) +
              (size_t)w * coop_mixed_wave_floats(A.m);
  const CovArgs32 c32 = cov32(c);
  const SiteRun run = xcd_site_run(A.coop_blocks, b, A.n, nw, w);
  // the wave's running sums live in LDS (lane 0 updates them), which keeps
  // this kernel's registers low enough for several waves per SIMD
  const int nv = 2 + A.nz * (A.nz + 1) / 2;
  for (int k = t; k < nv; k += 64) red[w][k] = 0.0;
  wave_sync();
  for (int i = run.first; i < run.end; i += run.stride) {
    double F, e[kMaxNZ], bcoef;
    int cnt, off;
    bool ok;
    if constexpr (MIXED)
      ok = coop_site_mixed<MM>(A, P, c, c32, tabs.bt32, Wf, i, F, e, bcoef,
                               cnt, off);
    else
      ok = coop_site<MM>(A, P, c, tabs.bt, tabs.mt, Am, i, F, e, bcoef, cnt,
                         off);
    if (!ok && t == 0) atomicMin(P.err, i);
    if (!BATCH) {
      if (A.B && ok && t < cnt) A.B[off + t] = bcoef;
      if (A.F && t == 0) A.F[i] = F;
    }
    if (t == 0) {
      const double rf = 1.0 / F;
      red[w][0] += log(F);
      int g = 2;
#pragma unroll
      for (int p = 0; p < kMaxNZ; p++) {
        if (p < A.nz) {
#pragma unroll
          for (int q = 0; q <= p; q++) red[w][g + q] += e[p] * e[q] * rf;
        }
        g += p + 1;
      }
    }
  }
  if (!P.rows) return;
  __syncthreads();
  const int k = (int)threadIdx.x;
  if (k < nv) {
    double s = 0.0;
    if (k != 1)
      for (int ww = 0; ww < nw; ww++) s += red[ww][k];
    P.rows[(size_t)k * A.rows_cap + b] = s;
  }
}

// ---------------------------------------------------------------------------
// one site per 16-lane row (four sites per wavefront), m <= 32
// ---------------------------------------------------------------------------
constexpr int kGrpSlots = 4 * kCoopMaxWaves;  // sites a block works on at once

// Per-wave LDS of the row kernel, in doubles: each of its four rows'
// neighbor coordinates, x then y.
__host__ __device__ inline int grp_wave_doubles(int mm) { return 4 * 2 * mm; }

// Covariance entries of the row kernel.  A lane evaluates about fifty per
// site in fully unrolled code, so only the common path is inline: for the
// matern (MAT) the pass's table, with near-coincident sites and x >= 64
// (outside the table) out of line; otherwise the exponential, spherical and
// gaussian.  The out-of-line series reads the parameters from the block's
// LDS copy `al` (a reference to the kernel argument would make the kernel
// keep a stack copy of it and read the inline path's fields from there).
__device__ __noinline__ double grp_matern_outside(double x, const CovArgs &al,
                                                  const BesselTab &bt) {
  if (x > 700.0) return 0.0;
  return al.mnorm * xnu_bessel_k(x, al, bt);
}

template <bool MAT>
__device__ __forceinline__ double grp_corr(double d2, const CovArgs &a,
                                           const CovArgs &al,
                                           const BesselTab &bt,
                                           const MaternTab &T) {
  if constexpr (MAT) {
    const double x = nn_sqrt(d2) * a.phi;
    if (!(x > 0.0)) return a.s2;
    const int o = matern_tab_octave(x);
    if ((unsigned)o >= (unsigned)kMatOct) return grp_matern_outside(x, al, bt);
    return matern_tab_at(x, o, T);
  } else {
    switch (a.cov) {
      case 0: return s2_corr<0>(d2, a);
      case 1: return s2_corr<1>(d2, a);
      default: return s2_corr<3>(d2, a);
    }
  }
}

// A site's inputs on its row: neighbor count and offset, the lane's two
// neighbor ids and coordinates, the site's coordinates.  The three loads
// depend on each other, so the row kernel fetches the next site's in three
// stages spread over the current site's work.
struct GrpIn {
  int cnt, off, ida, idb;
  double2 qa, qb, s;
};

__device__ __forceinline__ void grp_in_lu(const BfArgs &A, const BfProblem &P,
                                          int i, bool valid, GrpIn &in) {
  in.cnt = 0;
  in.off = 0;
  if (valid) {
    in.cnt = P.lu[A.n + i];
    in.off = P.lu[i];
  }
}

template <int R>
__device__ __forceinline__ void grp_in_ids(const BfProblem &P, int u,
                                           GrpIn &in) {
  in.ida = 0;
  in.idb = 0;
  if (u < in.cnt) in.ida = P.nnIndx[in.off + u];
  if (R == 2 && u + 16 < in.cnt) in.idb = P.nnIndx[in.off + u + 16];
}

template <int R>
__device__ __forceinline__ void grp_in_xy(const BfProblem &P, int u, int i,
                                          bool valid, GrpIn &in) {
  in.qa = make_double2(0.0, 0.0);
  in.qb = in.qa;
  in.s = in.qa;
  if (u < in.cnt) in.qa = P.xy[in.ida];
  if (R == 2 && u + 16 < in.cnt) in.qb = P.xy[in.idb];
  if (valid) in.s = P.xy[i];
}

// Site i on the 16-lane row of lane u (valid: i < n), its inputs in `in`.
// The lane builds rows u and u + 16 of the site's system in registers from
// the neighbor coordinates the row stages in px/py (rows past the site's
// count are identity rows with a zero right-hand side, so B and F are those
// of the cnt x cnt system), then factor, forward and back substitution
// (nngp_group.hpp).  Meanwhile it fetches the inputs of the row's next site
// in2 (site i2, valid2).  Every lane of the row returns F and the residuals
// e; false when a pivot or F is not positive.
template <int MM, bool MAT>
__device__ inline bool grp_site(const BfArgs &A, const BfProblem &P,
                                const CovArgs &c, const CovArgs &al,
                                const BesselTab &bt,
                                const MaternTab &mt, double *px, double *py,
                                int i, bool valid, const GrpIn &in, int i2,
                                bool valid2, GrpIn &in2, int u, double &F,
                                double (&e)[kMaxNZ]) {
  constexpr int R = GrpRows<MM>::R;
  // opaque per site: the compares of u with every row number below would
  // otherwise be hoisted out of the caller's site loop as lane masks, more
  // than the SGPR file holds
  u = lane_opaque(u);
  const int cnt = in.cnt, off = in.off, ida = in.ida, idb = in.idb;
  const double2 qa = in.qa, qb = in.qb, s = in.s;
  const bool ina = u < cnt, inb = R == 2 && u + 16 < cnt;
  grp_in_lu(A, P, i2, valid2, in2);  // next site, stage 1
  px[u] = qa.x;
  py[u] = qa.y;
  if constexpr (R == 2) {
    px[u + 16] = qb.x;
    py[u + 16] = qb.y;
  }
  wave_sync();
  GrpRows<MM> S;
  // row u: entries l < u, the diagonal, zeros above it
  static_for<0, 16>([&](auto lc) {
    constexpr int l = decltype(lc)::value;
    S.ra[l] = l == u ? (ina ? c.diag : 1.0) : 0.0;
    if (l < u && ina)
      S.ra[l] = grp_corr<MAT>(sqdist(qa.x, qa.y, px[l], py[l]), c, al, bt, mt);
  });
  if constexpr (R == 2) {  // row u + 16
    static_for<0, 32>([&](auto lc) {
      constexpr int l = decltype(lc)::value;
      S.rb[l] = l == u + 16 ? (inb ? c.diag : 1.0) : 0.0;
      if (l < u + 16 && inb)
        S.rb[l] = grp_corr<MAT>(sqdist(qb.x, qb.y, px[l], py[l]), c, al, bt, mt);
    });
  }
  const double ca =
      ina ? grp_corr<MAT>(sqdist(s.x, s.y, qa.x, qa.y), c, al, bt, mt) : 0.0;
  const double cb =
      inb ? grp_corr<MAT>(sqdist(s.x, s.y, qb.x, qb.y), c, al, bt, mt) : 0.0;
  wave_sync();  // the row's coordinates are read; the next site rewrites them
  grp_in_ids<R>(P, u, in2);  // next site, stage 2
  bool ok = grp_factor<MM>(S, u);
  grp_in_xy<R>(P, u, i2, valid2, in2);  // next site, stage 3
  double ya, yb, ba, bb;
  grp_forward<MM>(S, u, ca, cb, ya, yb);
  F = c.diag - row_sum16(ya * ya + yb * yb);
  grp_backward<MM>(S, u, ya, yb, ba, bb);
  ok = ok && F > 0.0;
  if (valid && A.B) {
    if (ina) A.B[off + u] = ba;
    if (inb) A.B[off + u + 16] = bb;
  }
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) {
    e[j] = 0.0;
    if (j < A.nz) {
      double part = 0.0;
      if (ina) part = ba * P.Z[(size_t)ida * A.zs + j];
      if (inb) part = fma(bb, P.Z[(size_t)idb * A.zs + j], part);
      part = row_sum16(part);
      e[j] = valid ? P.Z[(size_t)i * A.zs + j] - part : 0.0;
    }
  }
  if (!ok) {
    F = 1.0;  // keeps the sums finite; the pass reports the failure
#pragma unroll
    for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;
  }
  return ok;
}

// Logical block b (xcd_block) takes groups of four consecutive sites from
// its XCD's run (xcd_site_run over the groups); each wave works on one
// group, row g of the wave on site 4 q + g.  Each row's running sums live
// in its own LDS slot (lane 0 of the row updates them); the block's row of
// partial sums adds the slots in a fixed order.  Single problem, FP64.
// MAT: the matern (c.cov == 2), by the pass's table.
template <int MM, bool MAT>
__global__ __launch_bounds__(64 * kCoopMaxWaves) void k_bf_grp(BfArgs A) {
  extern __shared__ double smem[];
  __shared__ double red[kGrpSlots][kStatsMax];
  __shared__ CoopTabs<false> tabs;
  __shared__ CovArgs cl;  // for the out-of-line matern series
  const int nw = (int)blockDim.x >> 6;
  const int w = (int)threadIdx.x >> 6;
  const int t = (int)threadIdx.x & 63;
  const int g = t >> 4;
  const int u = t & 15;
  const int b = xcd_block(A.coop_blocks);
  if (b < 0) return;
  const BfProblem P = bf_problem<false>(A);
  const CovArgs c = A.c;
  if (MAT) {  // the block's Bessel-K divisors and the pass's table
    if (threadIdx.x == 0) cl = c;
    bessel_tab_fill(tabs.bt, c);
    __syncthreads();
    matern_tab_fill(tabs.mt, c, tabs.bt, smem);
    __syncthreads();  // its scratch becomes the waves' coordinate areas
  }
  double *px = smem + (size_t)w * grp_wave_doubles(MM) + (size_t)g * 2 * MM;
  double *py = px + MM;
  const int slot = 4 * w + g;
  const int nv = 2 + A.nz * (A.nz + 1) / 2;
  for (int k = u; k < nv; k += 16) red[slot][k] = 0.0;
  wave_sync();
  const SiteRun run = xcd_site_run(A.coop_blocks, b, (A.n + 3) >> 2, nw, w);
  GrpIn in;
  {  // the first site's inputs; later sites' arrive during their
     // predecessor's work
    const int i = 4 * run.first + g;
    const bool valid = run.first < run.end && i < A.n;
    grp_in_lu(A, P, i, valid, in);
    grp_in_ids<GrpRows<MM>::R>(P, u, in);
    grp_in_xy<GrpRows<MM>::R>(P, u, i, valid, in);
  }
  for (int q = run.first; q < run.end; q += run.stride) {
    const int i = 4 * q + g;
    const bool valid = i < A.n;
    const int q2 = q + run.stride, i2 = 4 * q2 + g;
    const bool valid2 = q2 < run.end && i2 < A.n;
    double F, e[kMaxNZ];
    GrpIn in2;
    const bool ok = grp_site<MM, MAT>(A, P, c, cl, tabs.bt, tabs.mt, px, py,
                                      i, valid, in, i2, valid2, in2, u, F, e);
    in = in2;
    if (valid && u == 0) {
      if (A.F) A.F[i] = F;
      if (!ok) atomicMin(P.err, i);
      const double rf = 1.0 / F;
      red[slot][0] += log(F);
      int gg = 2;
#pragma unroll
      for (int p = 0; p < kMaxNZ; p++) {
        if (p < A.nz) {
#pragma unroll
          for (int qq = 0; qq <= p; qq++)
            red[slot][gg + qq] += e[p] * e[qq] * rf;
        }
        gg += p + 1;
      }
    }
  }
  if (!P.rows) return;
  __syncthreads();
  const int k = (int)threadIdx.x;
  if (k < nv) {
    double sum = 0.0;
    if (k != 1)
      for (int sl = 0; sl < 4 * nw; sl++) sum += red[sl][k];
    P.rows[(size_t)k * A.rows_cap + b] = sum;
  }
}

// ---------------------------------------------------------------------------
// sums
// ---------------------------------------------------------------------------
// Block (k, j): out[j*out_stride + k] = sum over rows r < nrows of column k
// of problem j's rows (fixed per-thread order, fixed butterfly).  Block
// (0, j) also hands over problem j's failure flag:
// err_out[j] = err[j], err[j] = INT_MAX.
__global__ __launch_bounds__(kRowsThreads) void k_rows_sum(
    const double *rows, int cap, int nrows, long long ps_rows, double *out,
    int out_stride, int *err, double *err_out, SumsSignal sig) {
  __shared__ double red[kRowsThreads / 64];
  const int k = (int)blockIdx.x, j = (int)blockIdx.y;
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  const double *col = rows + (size_t)j * ps_rows + (size_t)k * cap;
  // eight independent partial sums, so that a thread's loads are in flight
  // together; the order (per thread, then the fixed combination) is fixed
  constexpr int U = 8;
  double acc[U];
#pragma unroll
  for (int u = 0; u < U; u++) acc[u] = 0.0;
  int r = (int)threadIdx.x;
  for (; r + (U - 1) * kRowsThreads < nrows; r += U * kRowsThreads) {
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] += col[r + u * kRowsThreads];
  }
#pragma unroll
  for (int u = 0; u < U; u++) {  // the tail: fewer than U rows left
    const int rr = r + u * kRowsThreads;
    if (rr < nrows) acc[u] += col[rr];
  }
  double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) +
             ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  s = wave_sum(s);
  if (t == 0) red[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
#pragma unroll
    for (int ww = 0; ww < kRowsThreads / 64; ww++) tot += red[ww];
    out[(size_t)j * out_stride + k] = tot;
    if (k == 0 && err) {
      err_out[j] = (double)err[j];
      err[j] = INT_MAX;
    }
    if (sig.flag) {  // the completion word, from the last block to finish
      __threadfence_system();  // this block's outputs, visible to the host
      const unsigned nb = gridDim.x * gridDim.y;
      if ((unsigned)atomicAdd(sig.done, 1) == nb - 1) {
        *sig.done = 0;
        __hip_atomic_store(sig.flag, sig.seq, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

int rows_sum_launch(const double *rows, int cap, int nrows, long long ps_rows,
                    int nv, int nprob, double *out, int out_stride, int *err,
                    double *err_out, hipStream_t st, const SumsSignal *sig) {
  const SumsSignal none;
  hipLaunchKernelGGL(k_rows_sum, dim3(nv, nprob), dim3(kRowsThreads), 0, st,
                     rows, cap, nrows, ps_rows, out, out_stride, err,
                     err_out, sig ? *sig : none);
  return (int)hipGetLastError();
}

static int red_grid(int n) {
  const int b = (n + kRedThreads - 1) / kRedThreads;
  return b < 1 ? 1 : (b > kRedBlocks ? kRedBlocks : b);
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
bool bf_lane_ok(int m, int variant) { return m <= 16 && variant >= 0; }

static int lane_M(int m);
static size_t mixed_copy_shm(int m, int threads);

// Scratch of matern_tab_fill, in doubles.
constexpr int kMatScratch = 2 * kMatPieces * kMatN + kMatN * kMatN;

// LDS per block of the wavefront kernel with `waves` waves (FP64; its
// system areas also hold the matern table's build).
static size_t coop_shm(int m, int waves) {
  const size_t d = (size_t)waves * coop_wave_doubles(m);
  return (d > (size_t)kMatScratch ? d : (size_t)kMatScratch) * sizeof(double);
}

bool bf_config_ok(const BfConfig &cfg, int m, int variant) {
  if (cfg.family == 0)
    return bf_lane_ok(m, variant) &&
           (cfg.threads == 64 || cfg.threads == 128 || cfg.threads == 256) &&
           (cfg.mode == 0 || cfg.mode == 2 || cfg.mode == 3 ||
            (cfg.mode == 1 && mixed_copy_shm(m, cfg.threads) <= 65536));
  if (cfg.family == 2)
    return m <= kGrpMaxM && cfg.mode == 0 &&
           (cfg.threads == 64 || cfg.threads == 128 || cfg.threads == 256);
  if (cfg.family != 1) return false;
  const int waves = cfg.threads / 64;
  return m <= kCoopMaxM && cfg.threads % 64 == 0 && waves >= 1 &&
         waves <= kCoopMaxWaves && coop_shm(m, waves) <= 65536;
}

// Untuned passes: per lane at 64 threads with the covariate rows staged
// in LDS where that applies (the fastest per-lane shape at n = 1e6, m = 15:
// 0.21 ms a pass against 0.35 ms at 256 threads without staging); else,
// for single FP64 problems with m <= 32, the row kernel at 256 threads
// (n = 4e6, m = 32, Bessel-K matern: 14.8 ms against 41 ms on the
// wavefront kernel); else the wavefront kernel.
BfConfig bf_default(int m, int variant, bool single_fp64) {
  BfConfig c;
  if (bf_lane_ok(m, variant)) {
    c.family = 0;
    c.threads = 64;
    c.mode = 2;
  } else if (single_fp64 && m <= kGrpMaxM) {
    c.family = 2;
    c.threads = 256;
  } else {
    c.family = 1;
    int w = kCoopMaxWaves;
    while (w > 1 && coop_shm(m, w) > 65536) w--;
    c.threads = 64 * w;
  }
  return c;
}

int bf_rows(const BfArgs &a, const BfConfig &cfg) {
  if (cfg.family == 0) return (a.n + cfg.threads - 1) / cfg.threads;
  const int waves = cfg.threads / 64;
  // the row kernel's waves take groups of four sites
  const int units = cfg.family == 2 ? (a.n + 3) / 4 : a.n;
  const int need = (units + waves - 1) / waves;
  return need < kCoopBlocks ? need : kCoopBlocks;
}

template <int M, bool BATCH>
static void launch_lane_m(const BfArgs &a, int variant, const dim3 &grid,
                          const dim3 &blk, hipStream_t st) {
  // LDS of the block sums: site_terms(nz) doubles a lane
  const size_t shm =
      a.rows ? (size_t)blk.x * site_terms(a.nz) * sizeof(double) : 0;
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_bf_lane<M, 0, BATCH>), grid, blk, shm, st, a); break;
    case 1: hipLaunchKernelGGL((k_bf_lane<M, 1, BATCH>), grid, blk, shm, st, a); break;
    case kCovMatern15: hipLaunchKernelGGL((k_bf_lane<M, kCovMatern15, BATCH>), grid, blk, shm, st, a); break;
    case kCovMatern25: hipLaunchKernelGGL((k_bf_lane<M, kCovMatern25, BATCH>), grid, blk, shm, st, a); break;
    default: hipLaunchKernelGGL((k_bf_lane<M, 3, BATCH>), grid, blk, shm, st, a); break;
  }
}

// Per-lane kernel size for m neighbors.
static int lane_M(int m) {
  return m <= 4 ? 4 : m <= 8 ? 8 : m <= 10 ? 10 : m <= 12 ? 12 : m <= 15 ? 15
                                                                          : 16;
}

// LDS of k_bf_lane_lds: per wavefront, M + 1 rows of zs doubles per lane.
static size_t lane_lds_shm(int m, int zs, int threads) {
  return (size_t)(threads / 64) * (lane_M(m) + 1) * (zs / 2) * 64 *
         sizeof(double2);
}

// Whether an FP64 pass can stage its covariate rows in LDS (mode 2: m > 8,
// where the factorization is long enough to cover the loads; batched
// passes with zs = 4 only); where it cannot, mode 2 runs the plain
// per-lane kernel, whose results are the same bit for bit.
static bool lane_lds_applies(const BfArgs &a, int threads) {
  return !a.mixed && a.Z && a.nz > 0 && a.m > 8 &&
         (a.zs == 4 || (a.zs == 2 && a.ng == 0)) &&
         ((uintptr_t)a.Z & 15) == 0 && (a.ng == 0 || a.fs_Z % 2 == 0) &&
         lane_lds_shm(a.m, a.zs, threads) <= 65536;
}

template <int M, int ZH, bool BATCH>
static void launch_lane_lds_mh(const BfArgs &a, int variant, const dim3 &grid,
                               const dim3 &blk, size_t shm, hipStream_t st) {
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_bf_lane_lds<M, 0, ZH, BATCH>), grid, blk, shm, st, a); break;
    case 1: hipLaunchKernelGGL((k_bf_lane_lds<M, 1, ZH, BATCH>), grid, blk, shm, st, a); break;
    case kCovMatern15: hipLaunchKernelGGL((k_bf_lane_lds<M, kCovMatern15, ZH, BATCH>), grid, blk, shm, st, a); break;
    case kCovMatern25: hipLaunchKernelGGL((k_bf_lane_lds<M, kCovMatern25, ZH, BATCH>), grid, blk, shm, st, a); break;
    default: hipLaunchKernelGGL((k_bf_lane_lds<M, 3, ZH, BATCH>), grid, blk, shm, st, a); break;
  }
}

template <int M>
static void launch_lane_lds_m(const BfArgs &a, int variant, const dim3 &grid,
                              const dim3 &blk, size_t shm, hipStream_t st) {
  if (a.ng > 0) launch_lane_lds_mh<M, 2, true>(a, variant, grid, blk, shm, st);
  else if (a.zs == 2) launch_lane_lds_mh<M, 1, false>(a, variant, grid, blk, shm, st);
  else launch_lane_lds_mh<M, 2, false>(a, variant, grid, blk, shm, st);
}

static void launch_lane_lds(const BfArgs &a, int variant, const dim3 &grid,
                            const dim3 &blk, hipStream_t st) {
  const size_t shm = lane_lds_shm(a.m, a.zs, (int)blk.x);
  switch (lane_M(a.m)) {
    case 10: launch_lane_lds_m<10>(a, variant, grid, blk, shm, st); break;
    case 12: launch_lane_lds_m<12>(a, variant, grid, blk, shm, st); break;
    case 15: launch_lane_lds_m<15>(a, variant, grid, blk, shm, st); break;
    default: launch_lane_lds_m<16>(a, variant, grid, blk, shm, st); break;
  }
}

// LDS of the mixed kernel's covariance copy: (M(M+1)/2 + M) floats a lane.
static size_t mixed_copy_shm(int m, int threads) {
  const int M = lane_M(m);
  return (size_t)threads * (M * (M + 1) / 2 + M) * sizeof(float);
}

// LDS of the mixed per-lane kernel without the covariance copy: the
// neighbor offsets.
static size_t mixed_offsets_shm(int m, int threads) {
  return (size_t)threads * 2 * lane_M(m) * sizeof(float);
}

template <int M, bool COPY>
static void launch_mixed_mc(const BfArgs &a, int variant, const dim3 &grid,
                            const dim3 &blk, size_t shm, hipStream_t st) {
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_bf_mixed<M, 0, COPY>), grid, blk, shm, st, a); break;
    case 1: hipLaunchKernelGGL((k_bf_mixed<M, 1, COPY>), grid, blk, shm, st, a); break;
    case kCovMatern15: hipLaunchKernelGGL((k_bf_mixed<M, kCovMatern15, COPY>), grid, blk, shm, st, a); break;
    case kCovMatern25: hipLaunchKernelGGL((k_bf_mixed<M, kCovMatern25, COPY>), grid, blk, shm, st, a); break;
    default: hipLaunchKernelGGL((k_bf_mixed<M, 3, COPY>), grid, blk, shm, st, a); break;
  }
}

// Rows of Z that k_bf_mixed_reg stages in LDS: ZH chunks of 16 bytes a
// row (zs = 2 ZH, Z 16-byte aligned, m > 8 as for the FP64 staging, and the
// block's stage within 64 KiB), else 0.
static int mixed_reg_zh(const BfArgs &a, int threads) {
  if (!a.Z || a.nz <= 0 || a.m <= 8 || ((uintptr_t)a.Z & 15) != 0) return 0;
  const int zh = a.zs == 2 ? 1 : a.zs == 4 ? 2 : 0;
  const size_t stage = (size_t)(threads / 64) *
                       mixed_reg_wave_chunks(lane_M(a.m), zh, a.nz) *
                       sizeof(double2);
  return stage <= 65536 ? zh : 0;
}

static size_t mixed_reg_shm(const BfArgs &a, int threads) {
  return (size_t)(threads / 64) *
         mixed_reg_wave_chunks(lane_M(a.m), mixed_reg_zh(a, threads), a.nz) *
         sizeof(double2);
}

template <int M, int ZH>
static void launch_mixed_reg_mh(const BfArgs &a, int variant, const dim3 &grid,
                                const dim3 &blk, size_t shm, hipStream_t st) {
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_bf_mixed_reg<M, 0, ZH>), grid, blk, shm, st, a); break;
    case 1: hipLaunchKernelGGL((k_bf_mixed_reg<M, 1, ZH>), grid, blk, shm, st, a); break;
    case kCovMatern15: hipLaunchKernelGGL((k_bf_mixed_reg<M, kCovMatern15, ZH>), grid, blk, shm, st, a); break;
    case kCovMatern25: hipLaunchKernelGGL((k_bf_mixed_reg<M, kCovMatern25, ZH>), grid, blk, shm, st, a); break;
    default: hipLaunchKernelGGL((k_bf_mixed_reg<M, 3, ZH>), grid, blk, shm, st, a); break;
  }
}

template <int M>
static void launch_mixed_reg(const BfArgs &a, int variant, const dim3 &grid,
                             const dim3 &blk, hipStream_t st) {
  const size_t shm = mixed_reg_shm(a, (int)blk.x);
  switch (mixed_reg_zh(a, (int)blk.x)) {
    case 1: launch_mixed_reg_mh<M, 1>(a, variant, grid, blk, shm, st); break;
    case 2: launch_mixed_reg_mh<M, 2>(a, variant, grid, blk, shm, st); break;
    default: launch_mixed_reg_mh<M, 0>(a, variant, grid, blk, shm, st); break;
  }
}

// mode 1: the covariance copy in LDS; 3: in registers (one wave per SIMD,
// covariate rows staged in LDS where they can be); else re-evaluated
template <int M>
static void launch_mixed_m(const BfArgs &a, int variant, int mode,
                           const dim3 &grid, const dim3 &blk, hipStream_t st) {
  if (mode == 1)
    launch_mixed_mc<M, true>(a, variant, grid, blk,
                             mixed_copy_shm(a.m, (int)blk.x), st);
  else if (mode == 3)
    launch_mixed_reg<M>(a, variant, grid, blk, st);
  else
    launch_mixed_mc<M, false>(a, variant, grid, blk,
                              mixed_offsets_shm(a.m, (int)blk.x), st);
}

static void launch_mixed(const BfArgs &a, int variant, int mode,
                         const dim3 &grid, const dim3 &blk, hipStream_t st) {
  switch (lane_M(a.m)) {
    case 4: launch_mixed_m<4>(a, variant, mode, grid, blk, st); break;
    case 8: launch_mixed_m<8>(a, variant, mode, grid, blk, st); break;
    case 10: launch_mixed_m<10>(a, variant, mode, grid, blk, st); break;
    case 12: launch_mixed_m<12>(a, variant, mode, grid, blk, st); break;
    case 15: launch_mixed_m<15>(a, variant, mode, grid, blk, st); break;
    default: launch_mixed_m<16>(a, variant, mode, grid, blk, st); break;
  }
}

template <bool BATCH>
static void launch_lane(const BfArgs &a, int variant, const dim3 &grid,
                        const dim3 &blk, hipStream_t st) {
  if (a.m <= 4) launch_lane_m<4, BATCH>(a, variant, grid, blk, st);
  else if (a.m <= 8) launch_lane_m<8, BATCH>(a, variant, grid, blk, st);
  else if (a.m <= 10) launch_lane_m<10, BATCH>(a, variant, grid, blk, st);
  else if (a.m <= 12) launch_lane_m<12, BATCH>(a, variant, grid, blk, st);
  else if (a.m <= 15) launch_lane_m<15, BATCH>(a, variant, grid, blk, st);
  else launch_lane_m<16, BATCH>(a, variant, grid, blk, st);
}

// The wavefront kernel's register-resident factor is sized by the smallest
// of 16, 32 and 64 rows that holds m.
template <bool BATCH, bool MIXED>
static void launch_coop(const BfArgs &a, const dim3 &grid, const dim3 &blk,
                        size_t shm, hipStream_t st) {
  switch (wave_rows_for(a.m)) {
    case 16:
      hipLaunchKernelGGL((k_bf_coop<BATCH, 16, MIXED>), grid, blk, shm, st, a);
      break;
    case 32:
      hipLaunchKernelGGL((k_bf_coop<BATCH, 32, MIXED>), grid, blk, shm, st, a);
      break;
    default:
      hipLaunchKernelGGL((k_bf_coop<BATCH, 64, MIXED>), grid, blk, shm, st, a);
  }
}

// LDS of the row kernel's block: the matern table's build, then the
// waves' coordinate areas.
static size_t grp_shm(const BfArgs &a, int waves, int mm) {
  const size_t d = (size_t)waves * grp_wave_doubles(mm);
  const size_t need =
      a.c.cov == 2 && d < (size_t)kMatScratch ? (size_t)kMatScratch : d;
  return need * sizeof(double);
}

static void launch_grp(const BfArgs &a, const dim3 &grid, const dim3 &blk,
                       hipStream_t st) {
  const int waves = (int)blk.x / 64;
  const int mm = a.m <= 16 ? 16 : 32;
  const size_t shm = grp_shm(a, waves, mm);
  const bool mat = a.c.cov == 2;
  if (mm == 16) {
    if (mat) hipLaunchKernelGGL((k_bf_grp<16, true>), grid, blk, shm, st, a);
    else hipLaunchKernelGGL((k_bf_grp<16, false>), grid, blk, shm, st, a);
  } else {
    if (mat) hipLaunchKernelGGL((k_bf_grp<32, true>), grid, blk, shm, st, a);
    else hipLaunchKernelGGL((k_bf_grp<32, false>), grid, blk, shm, st, a);
  }
}

int bf_launch(const BfArgs &a0, const BfConfig &cfg, int nprob,
              hipStream_t st) {
  if (a0.n <= 0 || nprob <= 0) return 0;
  const bool batch = a0.ng > 0;
  if (!batch && nprob != 1) return -2;
  const int variant = a0.variant;
  if (!bf_config_ok(cfg, a0.m, variant)) return -1;
  BfArgs a = a0;
  const int nrows = bf_rows(a, cfg);
  if (a.rows && nrows > a.rows_cap) return -2;
  const dim3 grid(8 * ((nrows + 7) / 8), nprob);
  const dim3 blk(cfg.threads);
  if (cfg.family == 0) {
    if (batch) {
      if (cfg.mode == 2 && lane_lds_applies(a, cfg.threads))
        launch_lane_lds(a, variant, grid, blk, st);
      else
        launch_lane<true>(a, variant, grid, blk, st);
    }
    else if (a.mixed) launch_mixed(a, variant, cfg.mode, grid, blk, st);
    else if (cfg.mode == 2 && lane_lds_applies(a, cfg.threads))
      launch_lane_lds(a, variant, grid, blk, st);
    else launch_lane<false>(a, variant, grid, blk, st);
  } else if (cfg.family == 2) {
    if (batch || a.mixed) return -1;  // single FP64 problems only
    a.coop_blocks = nrows;
    launch_grp(a, grid, blk, st);
  } else {
    a.coop_blocks = nrows;
    const int waves = cfg.threads / 64;
    if (batch)
      launch_coop<true, false>(a, grid, blk, coop_shm(a.m, waves), st);
    else if (a.mixed)
      launch_coop<false, true>(
          a, grid, blk,
          (size_t)waves * coop_mixed_wave_floats(a.m) * sizeof(float), st);
    else
      launch_coop<false, false>(a, grid, blk, coop_shm(a.m, waves), st);
  }
  return (int)hipGetLastError();
}

int bf_sums_launch(const BfArgs &a, const BfConfig &cfg, int nprob,
                   double *out, int out_stride, double *err_out,
                   hipStream_t st, const SumsSignal *sig) {
  const int nv = 2 + a.nz * (a.nz + 1) / 2;
  return rows_sum_launch(a.rows, a.rows_cap, bf_rows(a, cfg), a.ps_rows, nv,
                         nprob, out, out_stride, a.err, err_out, st, sig);
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
// Candidates: sites per wavefront (64: the per-lane kernel at 64, 128 or
// 256 threads per block; 1: the wavefront kernel with 1, 2 or 4 waves per
// block, i.e. LDS tiles of 1, 2 or 4 systems).  Each is timed with HIP
// events on the caller's stream (one warm-up, then `reps` launches) and the
// fastest kept.  Results are cached per (m, variant, nz, size class).
namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, BfConfig> g_tune_cache;
}  // namespace

int bf_autotune(const BfArgs &a, int nprob, hipStream_t st, BfConfig *best,
                BfTuneReport *rep) {
  const int variant = a.variant;
  const long long work = (long long)a.n * nprob;
  int cls = 0;
  while ((1LL << (cls + 1)) <= work) cls++;
  const auto key = std::make_tuple(a.m, variant * 4 + a.mixed, a.nz, cls);
  if (!rep) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *best = it->second;
      return 0;
    }
  }
  std::vector<BfConfig> cand;
  if (a.mixed) {  // the per-lane mixed kernel: block size x covariance
                  // copy (none, LDS, registers); then the wavefront one
    for (int th : {64, 128, 256})
      for (int mode : {0, 1, 3}) cand.push_back(BfConfig{0, th, mode});
  } else {  // the per-lane kernel: block size x covariate rows in LDS
    for (int th : {64, 128, 256}) {
      cand.push_back(BfConfig{0, th, 0});
      if (lane_lds_applies(a, th)) cand.push_back(BfConfig{0, th, 2});
    }
  }
  for (int w : {1, 2, 4}) cand.push_back(BfConfig{1, 64 * w, 0});
  if (!a.mixed && a.ng == 0)  // four sites per wavefront, one per row
    for (int th : {64, 128, 256}) cand.push_back(BfConfig{2, th, 0});
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  const int reps = 3;
  double best_ms = 1e300;
  BfConfig pick = bf_default(a.m, variant, !a.mixed && a.ng == 0);
  int nc = 0;
  int rc = 0;
  for (const BfConfig &cfg : cand) {
    if (!bf_config_ok(cfg, a.m, variant)) continue;
    if (a.rows && bf_rows(a, cfg) > a.rows_cap) continue;
    if ((rc = bf_launch(a, cfg, nprob, st)) != 0) break;
    if ((rc = (int)hipEventRecord(e0, st)) != 0) break;
    for (int r = 0; r < reps && rc == 0; r++) rc = bf_launch(a, cfg, nprob, st);
    if (rc) break;
    if ((rc = (int)hipEventRecord(e1, st)) != 0) break;
    if ((rc = (int)hipEventSynchronize(e1)) != 0) break;
    float ms = 0.0f;
    if ((rc = (int)hipEventElapsedTime(&ms, e0, e1)) != 0) break;
    const double per = ms / reps;
    if (rep && nc < NNGP_TUNE_MAX) {
      rep->family[nc] = cfg.family;
      rep->threads[nc] = cfg.threads;
      rep->mode[nc] = cfg.mode;
      rep->ms[nc] = per;
    }
    nc++;
    if (per < best_ms) {
      best_ms = per;
      pick = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (rep) {
    rep->n = nc < NNGP_TUNE_MAX ? nc : NNGP_TUNE_MAX;
    rep->best = pick;
    rep->best_ms = best_ms;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = pick;
  *best = pick;
  return 0;
}

// ---------------------------------------------------------------------------
// plain Q(u, v)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kRedThreads) void k_quadform(QArgs A) {
  __shared__ double red[kRedThreads / 64];
  double acc = 0.0;
  const int stride = (int)gridDim.x * kRedThreads;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < A.n; i += stride) {
    const int cnt = A.lu[A.n + i], off = A.lu[i];
    double a = 0.0, b = 0.0;
    for (int j = 0; j < cnt; j++) {
      const double bj = A.B[off + j];
      const int id = A.nnIndx[off + j];
      a += bj * A.u[id];
      b += bj * A.v[id];
    }
    acc += (A.u[i] - a) * (A.v[i] - b) / A.F[i];
  }
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  const double x = wave_sum(acc);
  if (t == 0) red[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < kRedThreads / 64; ww++) s += red[ww];
    A.rows[blockIdx.x] = s;
  }
}

static_assert(kRedBlocks == kQPartialRows,
              "k_quadform's grid bounds its partial sums");

int q_partial_rows(int n) { return red_grid(n); }

int q_launch(const QArgs &a, hipStream_t st) {
  const int nb = red_grid(a.n);
  hipLaunchKernelGGL(k_quadform, dim3(nb), dim3(kRedThreads), 0, st, a);
  return rows_sum_launch(a.rows, kRedBlocks, nb, 0, 1, 1, a.out, 1, nullptr,
                         nullptr, st);
}

__global__ void k_interleave(const double *cm, int n, double2 *xy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xy[i] = make_double2(cm[i], cm[(size_t)n + i]);
}

int interleave_launch(const double *coords_cm, int n, double2 *xy,
                      hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_interleave, dim3((n + 255) / 256), dim3(256), 0, st,
                     coords_cm, n, xy);
  return (int)hipGetLastError();
}

}  // namespace nngp
