// nngp_bf.hip — neighbor-system kernels: updateBF / updateConjBF
// (src/rNNGP.cpp:46-101, src/cNNGP.cpp:24-69), the quadratic form Q
// (src/util.cpp:261-280) and the site sums the samplers need.
//
// Two updateBF kernels:
//  * k_bf_lane<M, COV>: one site per lane, the whole M x M system in
//    registers, fully unrolled (M <= 16, closed-form covariances).  Sites
//    with fewer than M neighbors (the first m sites of the NNGP order) are
//    padded with far-away points: their rows/columns of C are exactly zero
//    off the diagonal, so the real block factors bit-for-bit as if unpadded
//    and the padded coefficients come out exactly 0.  The kernel also adds
//    up its block's site terms (log F_i and e_i e_i^T / F_i, with e_i the
//    residual vector of X and y at site i), so one launch gives everything
//    a Metropolis step needs.
//  * k_bf_coop: one site per wavefront, the matrix in LDS, one row per lane
//    (any m <= 64, any covariance including the Bessel-K matern); it stores
//    the residual vectors and k_stats adds them up.
// Both compute C = s2 R(N,N) + t2 I, c = s2 R(N,i), the Cholesky factor of C
// (pivots kept as reciprocal square roots), v = L^-1 c, B = L^-T v and
// F = s2 + t2 - v.v, i.e. the reference's potrf -> potri -> symv result
// without forming the inverse.
//
// Sums are deterministic: a fixed grid, every thread adds a fixed set of
// sites in a fixed order, each block reduces with a fixed butterfly per
// wavefront and its wavefronts in order into one partial row, and
// k_rows_sum adds the rows of each column in a fixed order.
#include <climits>

#include "nngp_kernels.hpp"

namespace nngp {

constexpr int kLaneBlock = 256;
constexpr int kCoopWaves = 4;
constexpr int kCoopMaxM = 64;
constexpr int kRedThreads = 256;  // threads per block of the sum kernels
constexpr int kRedBlocks = 1024;  // grid of the sum kernels (at most)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// DPP lane move of a double (two 32-bit moves); lanes outside ROWMASK get 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL,
                                             ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL,
                                             ROWMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes with DPP moves only (no LDS round trips); the
// total is valid in lane 63.  Fixed order: pairs, quads, half rows, rows,
// then rows 0+1 and 2+3, then all four.
__device__ __forceinline__ double wave_sum_to63(double x) {
  x += dpp_mov<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
  x += dpp_mov<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
  x += dpp_mov<0x141, 0xF>(x);  // row_half_mirror
  x += dpp_mov<0x140, 0xF>(x);  // row_mirror
  x += dpp_mov<0x142, 0xA>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_mov<0x143, 0xC>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

int rows_capacity(int n) {
  const int lane_rows = (n + kLaneBlock - 1) / kLaneBlock;
  return lane_rows > kRedBlocks ? lane_rows : kRedBlocks;
}

// ---------------------------------------------------------------------------
// one site per lane
// ---------------------------------------------------------------------------

// The block's partial row of one pass: column 0 sum log F_i, column 1 zero
// (quadratic forms are formed from the cross-products), then the lower
// triangle of e e^T / F row by row.  Called by every thread of the block.
template <int NT>
__device__ inline void block_site_terms(const BfArgs &A, bool valid,
                                        double F, const double (&e)[kMaxNZ],
                                        int row) {
  __shared__ double red[NT / 64][kStatsMax];
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  const double rf = valid ? 1.0 / F : 0.0;
  double x = wave_sum_to63(valid ? log(F) : 0.0);
  if (t == 63) red[w][0] = x;
  int g = 2;
#pragma unroll
  for (int p = 0; p < kMaxNZ; p++) {
    if (p < A.nz) {
#pragma unroll
      for (int q = 0; q <= p; q++) {
        x = wave_sum_to63(e[p] * e[q] * rf);
        if (t == 63) red[w][g + q] = x;
      }
    }
    g += p + 1;
  }
  __syncthreads();
  const int nv = 2 + A.nz * (A.nz + 1) / 2;
  const int k = (int)threadIdx.x;
  if (k < nv) {
    double s = 0.0;
    if (k != 1)
#pragma unroll
      for (int ww = 0; ww < NT / 64; ww++) s += red[ww][k];
    A.rows[(size_t)k * A.rows_cap + row] = s;
  }
}

template <int M, int COV>
__global__ __launch_bounds__(kLaneBlock) void k_bf_lane(BfArgs A) {
  const int b = xcd_block((A.n + kLaneBlock - 1) / kLaneBlock);
  if (b < 0) return;  // whole padding block
  const int i = b * kLaneBlock + (int)threadIdx.x;
  const bool valid = i < A.n;
  const CovArgs &c = A.c;
  double Fi = 1.0;
  double e[kMaxNZ];
#pragma unroll
  for (int j = 0; j < kMaxNZ; j++) e[j] = 0.0;

  if (valid) {
    const int cnt = A.lu[A.n + i];
    const int off = A.lu[i];
    if (cnt == 0) {  // site 0: F = s2 + t2, residual = the data
      Fi = c.diag;
#pragma unroll
      for (int j = 0; j < kMaxNZ; j++)
        if (j < A.nz) e[j] = A.Z[(size_t)i * A.zs + j];
    } else {
      // The index array carries 64 padding entries, so off + k is in
      // bounds; entries past this site's count are not its neighbors and
      // are replaced by index 0 before any use.
      int nb[M];
#pragma unroll
      for (int k = 0; k < M; k++) nb[k] = A.nnIndx[off + k];
      if (cnt < M) {
#pragma unroll
        for (int k = 0; k < M; k++)
          if (k >= cnt) nb[k] = 0;
      }
      double px[M], py[M];
#pragma unroll
      for (int k = 0; k < M; k++) {
        const double2 q = A.xy[nb[k]];
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
      const double2 s = A.xy[i];

      constexpr int T = M * (M + 1) / 2;
      double a[T];  // packed lower triangle
      double v[M];
#pragma unroll
      for (int k = 0; k < M; k++) {
        v[k] = s2_corr<COV>(sqdist(s.x, s.y, px[k], py[k]), c);
#pragma unroll
        for (int l = 0; l < k; l++)
          a[tri(k, l)] = s2_corr<COV>(sqdist(px[k], py[k], px[l], py[l]), c);
        a[tri(k, k)] = c.diag;
      }
      double vv;
      const bool ok = chol_solve<M>(a, v, vv);  // v <- B_i = C^-1 c
      if (A.B) {
#pragma unroll
        for (int k = 0; k < M; k++)
          if (k < cnt) A.B[off + k] = v[k];
      }
      Fi = c.diag - vv;
      // residual vectors; padded coefficients are exactly 0
#pragma unroll
      for (int j = 0; j < kMaxNZ; j++) {
        if (j < A.nz) {
          double ej = A.Z[(size_t)i * A.zs + j];
#pragma unroll
          for (int k = 0; k < M; k++)
            ej = fma(-v[k], A.Z[(size_t)nb[k] * A.zs + j], ej);
          e[j] = ej;
        }
      }
      if (!ok) atomicMin(A.err, i);
    }
    A.F[i] = Fi;
    if (A.E) {
#pragma unroll
      for (int j = 0; j < kMaxNZ; j++)
        if (j < A.nz) A.E[(size_t)j * A.n + i] = e[j];
    }
  }
  if (A.rows) block_site_terms<kLaneBlock>(A, valid, Fi, e, b);
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
// Waves per block so that the block's LDS stays within 64 KB.
static int coop_waves(int m) {
  const int per = coop_wave_doubles(m) * (int)sizeof(double);
  int w = 65536 / per;
  return w < 1 ? 1 : (w > kCoopWaves ? kCoopWaves : w);
}

__global__ __launch_bounds__(64 * kCoopWaves) void k_bf_coop(BfArgs A) {
  extern __shared__ double smem[];
  const int nw = (int)blockDim.x >> 6;
  const int ld = A.m + 1;
  const int per_wave = coop_wave_doubles(A.m);
  const int w = (int)threadIdx.x >> 6;
  const int t = (int)threadIdx.x & 63;
  double *Am = smem + (size_t)w * per_wave;  // row-major, Am[k*ld + l]
  double *cv = Am + A.m * ld;
  double *px = cv + A.m;
  double *py = px + A.m;

  const int nlog = (A.n + nw - 1) / nw;
  const int b = xcd_block(nlog);
  if (b < 0) return;
  const int i = b * nw + w;
  if (i >= A.n) return;  // whole wave leaves; no block-level barrier below
  const CovArgs &c = A.c;
  const int cnt = A.lu[A.n + i];
  if (cnt == 0) {
    if (t == 0) {
      A.F[i] = c.diag;
      if (A.E)
        for (int j = 0; j < A.nz; j++)
          A.E[(size_t)j * A.n + i] = A.Z[(size_t)i * A.zs + j];
    }
    return;
  }
  const int off = A.lu[i];
  const double2 s = A.xy[i];
  int myid = 0;
  if (t < cnt) {
    myid = A.nnIndx[off + t];
    const double2 q = A.xy[myid];
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
      const double val = s2_corr_rt(sqdist(ax, ay, px[l], py[l]), c);
      if (k < cnt)
        Am[k * ld + l] = val;
      else
        cv[l] = val;
    }
  }
  wave_sync();
  // right-looking Cholesky, lane t owns row t; diagonal keeps 1/L_jj
  bool ok = true;
  for (int j = 0; j < cnt; j++) {
    const double piv = Am[j * ld + j];
    ok = ok && (piv > 0.0);
    if (!ok) break;
    const double r = nn_rsqrt(piv);
    double ltj = 0.0;
    if (t > j && t < cnt) {
      ltj = Am[t * ld + j] * r;
      Am[t * ld + j] = ltj;
    }
    wave_sync();
    if (t == j) Am[j * ld + j] = r;
    if (t > j && t < cnt) {
      for (int l = j + 1; l <= t; l++)
        Am[t * ld + l] = fma(-ltj, Am[l * ld + j], Am[t * ld + l]);
    }
    wave_sync();
  }
  if (!ok) {
    if (t == 0) {
      atomicMin(A.err, i);
      A.F[i] = 1.0;  // keeps the sums finite; the pass reports the failure
      if (A.E)
        for (int j = 0; j < A.nz; j++) A.E[(size_t)j * A.n + i] = 0.0;
    }
    return;
  }
  // forward solve: lane t holds x_t
  double x = t < cnt ? cv[t] : 0.0;
  for (int k = 0; k < cnt; k++) {
    const double vk = readlane_d(x, k) * Am[k * ld + k];
    if (t == k) x = vk;
    if (t > k && t < cnt) x = fma(-Am[t * ld + k], vk, x);
  }
  const double vv = wave_sum(t < cnt ? x * x : 0.0);
  // back solve B = L^-T v
  for (int k = cnt - 1; k >= 0; k--) {
    const double bk = readlane_d(x, k) * Am[k * ld + k];
    if (t == k) x = bk;
    if (t < k) x = fma(-Am[k * ld + t], bk, x);
  }
  if (A.B && t < cnt) A.B[off + t] = x;
  if (t == 0) A.F[i] = c.diag - vv;
  if (A.E) {
    for (int j = 0; j < A.nz; j++) {
      const double part =
          wave_sum(t < cnt ? x * A.Z[(size_t)myid * A.zs + j] : 0.0);
      if (t == 0)
        A.E[(size_t)j * A.n + i] = A.Z[(size_t)i * A.zs + j] - part;
    }
  }
}

// ---------------------------------------------------------------------------
// sums
// ---------------------------------------------------------------------------
// out[k] = sum over rows r < nrows of rows[k*cap + r] (one block per column
// k, fixed per-thread order, fixed butterfly); block 0 also hands over the
// updateBF failure flag: *err_out = *err, *err = INT_MAX.
__global__ __launch_bounds__(kRedThreads) void k_rows_sum(
    const double *rows, int cap, int nrows, double *out, int *err,
    double *err_out) {
  __shared__ double red[kRedThreads / 64];
  const int k = (int)blockIdx.x;
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
  double s = 0.0;
  const double *col = rows + (size_t)k * cap;
  for (int r = (int)threadIdx.x; r < nrows; r += kRedThreads) s += col[r];
  s = wave_sum(s);
  if (t == 0) red[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
#pragma unroll
    for (int ww = 0; ww < kRedThreads / 64; ww++) tot += red[ww];
    out[k] = tot;
    if (k == 0 && err) {
      *err_out = (double)*err;
      *err = INT_MAX;
    }
  }
}

static int rows_sum_launch(const double *rows, int cap, int nrows, int nv,
                           double *out, int *err, double *err_out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_rows_sum, dim3(nv), dim3(kRedThreads), 0, st, rows,
                     cap, nrows, out, err, err_out);
  return (int)hipGetLastError();
}

static int red_grid(int n) {
  const int b = (n + kRedThreads - 1) / kRedThreads;
  return b < 1 ? 1 : (b > kRedBlocks ? kRedBlocks : b);
}

template <int NZ>
__global__ __launch_bounds__(kRedThreads) void k_stats(StatsArgs A) {
  constexpr int NV = 2 + NZ * (NZ + 1) / 2;
  __shared__ double red[kRedThreads / 64][NV];
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0.0;
  const int stride = (int)gridDim.x * kRedThreads;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < A.n; i += stride) {
    const double f = A.F[i];
    acc[0] += log(f);
    if constexpr (NZ > 0) {
      const double rf = 1.0 / f;
      double e[NZ];
#pragma unroll
      for (int j = 0; j < NZ; j++) e[j] = A.E[(size_t)j * A.n + i];
      int g = 2;
#pragma unroll
      for (int p = 0; p < NZ; p++)
#pragma unroll
        for (int q = 0; q <= p; q++) acc[g++] += e[p] * e[q] * rf;
    }
  }
  const int w = (int)threadIdx.x >> 6, t = (int)threadIdx.x & 63;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double x = wave_sum(acc[v]);
    if (t == 0) red[w][v] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < kRedThreads / 64; ww++) s += red[ww][threadIdx.x];
    A.rows[(size_t)threadIdx.x * A.rows_cap + blockIdx.x] = s;
  }
}

int stats_launch(const StatsArgs &a, hipStream_t st) {
  const int nb = red_grid(a.n);
  const int nz = a.E ? a.nz : 0;
  switch (nz) {
#define NNGP_STATS_CASE(K)                                                  \
  case K:                                                                  \
    hipLaunchKernelGGL(k_stats<K>, dim3(nb), dim3(kRedThreads), 0, st, a); \
    break;
    NNGP_STATS_CASE(0)
    NNGP_STATS_CASE(2)
    NNGP_STATS_CASE(3)
    NNGP_STATS_CASE(4)
    NNGP_STATS_CASE(5)
    NNGP_STATS_CASE(6)
    NNGP_STATS_CASE(7)
    NNGP_STATS_CASE(8)
    NNGP_STATS_CASE(9)
#undef NNGP_STATS_CASE
    default: return -1;
  }
  const int nv = 2 + nz * (nz + 1) / 2;
  return rows_sum_launch(a.rows, a.rows_cap, nb, nv, a.out, a.err,
                         a.err_out, st);
}

// ---------------------------------------------------------------------------
// dispatch of one pass
// ---------------------------------------------------------------------------
template <int M>
static void launch_lane_m(const BfArgs &a, int cov, hipStream_t st) {
  const int nlog = (a.n + kLaneBlock - 1) / kLaneBlock;
  const dim3 grid(8 * ((nlog + 7) / 8));
  switch (cov) {
    case 0: hipLaunchKernelGGL((k_bf_lane<M, 0>), grid, dim3(kLaneBlock), 0, st, a); break;
    case 1: hipLaunchKernelGGL((k_bf_lane<M, 1>), grid, dim3(kLaneBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_bf_lane<M, 3>), grid, dim3(kLaneBlock), 0, st, a); break;
  }
}

bool bf_fuses_sums(int m, int cov) { return cov != 2 && m <= 16; }

int bf_launch(const BfArgs &a, hipStream_t st) {
  if (a.n <= 0) return 0;
  const int cov = a.c.cov;
  if (bf_fuses_sums(a.m, cov)) {
    if (a.m <= 4) launch_lane_m<4>(a, cov, st);
    else if (a.m <= 8) launch_lane_m<8>(a, cov, st);
    else if (a.m <= 10) launch_lane_m<10>(a, cov, st);
    else if (a.m <= 12) launch_lane_m<12>(a, cov, st);
    else if (a.m <= 15) launch_lane_m<15>(a, cov, st);
    else launch_lane_m<16>(a, cov, st);
  } else {
    if (a.m > kCoopMaxM) return -1;
    const int nw = coop_waves(a.m);
    const int nlog = (a.n + nw - 1) / nw;
    const dim3 grid(8 * ((nlog + 7) / 8));
    const size_t shm = (size_t)nw * coop_wave_doubles(a.m) * sizeof(double);
    hipLaunchKernelGGL(k_bf_coop, grid, dim3(64 * nw), shm, st, a);
  }
  return (int)hipGetLastError();
}

int bf_sums_launch(const BfArgs &a, double *out, double *err_out,
                   hipStream_t st) {
  const int nv = 2 + a.nz * (a.nz + 1) / 2;
  const int nrows = (a.n + kLaneBlock - 1) / kLaneBlock;
  return rows_sum_launch(a.rows, a.rows_cap, nrows, nv, out, a.err, err_out,
                         st);
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

int q_launch(const QArgs &a, hipStream_t st) {
  const int nb = red_grid(a.n);
  hipLaunchKernelGGL(k_quadform, dim3(nb), dim3(kRedThreads), 0, st, a);
  return rows_sum_launch(a.rows, kRedBlocks, nb, 1, a.out, nullptr, nullptr,
                         st);
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
