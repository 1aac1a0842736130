// nngp_knn.hip — exact nearest-neighbor search on gfx950.
//
// mkNNIndxCB (src/nn.cpp:128-257) finds, for every site i, the m nearest
// sites among its NNGP predecessors (index < i), ascending by the squared
// distance fl(fl(dx*dx) + fl(dy*dy)), then stores sqrt of it; mkNNIndx
// (src/nn.cpp:63-111) does the same by a scan, comparing the Euclidean
// distances.  The reference walks a u = x + y sorted list; this search uses
// a strip index (points sorted by x, cut into strips, each strip sorted by
// y) and one query per lane.  Every pruning test compares a rounded partial
// square (dx^2 of the strip gap, or dy^2 inside a strip) with the current
// worst distance; since fl(dx^2) + fl(dy^2) rounds monotonically, nothing
// that could enter the list is ever skipped, so the result is the exact set
// of smallest distances in the reference's own arithmetic.  Products and
// sums here must not contract into FMAs (the reference rounds dx*dx, dy*dy
// and their sum separately), so contraction is switched off for this file.
//
// Ties.  Where no two of a site's m+1 smallest distances are equal, every
// correct search gives the same list.  Where some are, the reference's
// result depends on the order its walk meets the candidates and on its
// unstable shell sort, and near-equal squared distances can also swap
// through the rounding of the walk's pruning test (du)^2 > 2 * worst.  The
// strip search therefore keeps m+1 candidates and flags a site when two
// consecutive ones are closer than that rounding can reach (code book) or
// have equal Euclidean distances (scan); the flagged sites are compacted
// and redone by k_knn_replay in the reference's own candidate order
// (nngp_knn_replay.hpp).  Continuous coordinates almost never flag a site;
// lattices, rounded coordinates and duplicated sites flag most of them.
//
// nn2 (RANN) orders equal distances by ascending fitted row: the candidate
// key is the pair (squared distance, row).
#include <hipcub/hipcub.hpp>

#include <vector>

#include "nngp_host.hpp"
#include "nngp_knn_replay.hpp"
#include "nngp_strips.hpp"

#pragma clang fp contract(off)

namespace nngp {

constexpr int kKnnBlock = 256;

template <int M>
struct TopM {
  // Sorted ascending, right-aligned: slots [0, M - cnt) hold -inf and never
  // move, so d[M-1] is always the current cnt-th smallest distance.
  double d[M];
  int id[M];
  __device__ __forceinline__ void init(int cnt) {
#pragma unroll
    for (int k = 0; k < M; k++) {
      d[k] = k < M - cnt ? -__builtin_inf() : __builtin_inf();
      id[k] = -1;
    }
  }
  __device__ __forceinline__ double worst() const { return d[M - 1]; }
  __device__ __forceinline__ int worst_id() const { return id[M - 1]; }
  // (distance, id) pairs ordered lexicographically: equal distances by id
  __device__ __forceinline__ void insert_lex(double x, int ix) {
#pragma unroll
    for (int k = M - 1; k >= 1; k--) {
      const bool sh = d[k - 1] > x || (d[k - 1] == x && id[k - 1] > ix);
      const bool here = d[k] > x || (d[k] == x && id[k] > ix);
      const double nd = sh ? d[k - 1] : (here ? x : d[k]);
      const int ni = sh ? id[k - 1] : (here ? ix : id[k]);
      d[k] = nd;
      id[k] = ni;
    }
    if (d[0] > x || (d[0] == x && id[0] > ix)) {
      d[0] = x;
      id[0] = ix;
    }
  }
  __device__ __forceinline__ void insert(double x, int ix) {
#pragma unroll
    for (int k = M - 1; k >= 1; k--) {
      const bool sh = d[k - 1] > x;
      const bool here = d[k] > x;
      const double nd = sh ? d[k - 1] : (here ? x : d[k]);
      const int ni = sh ? id[k - 1] : (here ? ix : id[k]);
      d[k] = nd;
      id[k] = ni;
    }
    if (d[0] > x) {
      d[0] = x;
      id[0] = ix;
    }
  }
};

__device__ __forceinline__ int lower_bound_y(const double *y, int len,
                                             double v) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (y[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Scan strip t; `self` >= 0 restricts candidates to ids < self.  LEX keys
// candidates by (squared distance, id), else by the distance alone with
// strict insertion.
template <int M, bool LEX>
__device__ inline void scan_strip(const StripIndex &S, int t, double qx,
                                  double qy, int self, TopM<M> &top) {
  const int base = t * S.S;
  const int len = min(S.S, S.n - base);
  const double *sy = S.sy + base;
  const double *sx = S.sx + base;
  const int *sid = S.sid + base;
  const int p = lower_bound_y(sy, len, qy);
  // candidate j: squared distance rounded as the reference rounds it,
  // (x0-c0)^2 + (x1-c1)^2 with each square and the sum rounded; returns
  // false once dy^2 alone exceeds the current worst distance
  auto visit = [&](int j) {
    const double dy = qy - sy[j];
    const double dy2 = dy * dy;
    if (dy2 > top.worst()) return false;
    const int id = sid[j];
    if (self < 0 || id < self) {
      const double dx = qx - sx[j];
      const double dx2 = dx * dx;
      const double d2 = dx2 + dy2;
      if (LEX) {
        if (d2 < top.worst() || (d2 == top.worst() && id < top.worst_id()))
          top.insert_lex(d2, id);
      } else if (d2 < top.worst()) {
        top.insert(d2, id);
      }
    }
    return true;
  };
  for (int j = p; j < len; j++)
    if (!visit(j)) break;
  for (int j = p - 1; j >= 0; j--)
    if (!visit(j)) break;
}

template <int M, bool LEX>
__device__ inline void search(const StripIndex &S, int t0, double qx,
                              double qy, int self, bool left_only,
                              TopM<M> &top) {
  scan_strip<M, LEX>(S, t0, qx, qy, self, top);
  for (int t = t0 - 1; t >= 0; t--) {
    const double gap = qx - S.xhi[t];
    if (gap > 0.0 && gap * gap > top.worst()) break;
    scan_strip<M, LEX>(S, t, qx, qy, self, top);
  }
  if (left_only) return;
  for (int t = t0 + 1; t < S.nstrip; t++) {
    const double gap = S.xlo[t] - qx;
    if (gap > 0.0 && gap * gap > top.worst()) break;
    scan_strip<M, LEX>(S, t, qx, qy, self, top);
  }
}

// src/util.cpp:55-70 layout
__device__ __forceinline__ void nn_layout(int i, int m, int &off, int &cnt) {
  if (i == 0) {
    off = 0;
    cnt = 0;
  } else if (i < m) {
    off = (int)((double)i / 2 * (i - 1));
    cnt = i;
  } else {
    off = (int)((double)m / 2 * (m - 1) + (double)(i - m) * m);
    cnt = m;
  }
}

// One lane per site.  M >= m + 1 holds the cs = min(i, m + 1) smallest
// squared distances among the site's predecessors; the first cnt are the
// site's list, and all cs decide whether the site is flagged for replay:
//   code book (mode 0): two consecutive squared distances within
//     g = 2^-40 d + 2 sqrt(2 d) E + E^2 (d the largest of them).  The
//     reference's walk drops a direction when fl(du)^2 > 2 * worst; fl(du)
//     carries up to E = 2^-48 max(|x| + |y|) of absolute rounding, so a
//     candidate can be dropped, or a tie decided, against the exact order
//     only when its squared distance is within g of another's;
//   scan (mode 1): two consecutive Euclidean distances equal.
template <int M>
__global__ __launch_bounds__(kKnnBlock) void k_knn_pred(KnnArgs A) {
  const int i = blockIdx.x * kKnnBlock + (int)threadIdx.x;
  if (i >= A.n) return;
  int off, cnt;
  nn_layout(i, A.m, off, cnt);
  A.lu[i] = off;
  A.lu[A.n + i] = cnt;
  if (cnt == 0) {
    A.flag[i] = 0;
    return;
  }
  const int cs = min(i, A.m + 1);
  const double2 q = A.xy[i];
  TopM<M> top;
  top.init(cs);
  search<M, false>(A.s, A.strip_of[i], q.x, q.y, i, A.sorted != 0, top);
  // slots below M - cs hold -inf: their differences are NaN or +inf and
  // their square roots NaN, so they never flag
  bool flag = false;
  if (A.mode == 0) {
    const double dmax = top.d[M - 1];
    const double g =
        ldexp(dmax, -40) + 2.0 * sqrt(2.0 * dmax) * A.E + A.E * A.E;
#pragma unroll
    for (int k = 0; k + 1 < M; k++) flag |= top.d[k + 1] - top.d[k] <= g;
  } else {
#pragma unroll
    for (int k = 0; k + 1 < M; k++)
      flag |= sqrt(top.d[k + 1]) == sqrt(top.d[k]);
  }
  A.flag[i] = flag ? 1 : 0;
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = k - (M - cs);  // rank among the site's candidates
    if (r >= 0 && r < cnt) {
      A.nnIndx[off + r] = top.id[k];
      if (A.nnDist) A.nnDist[off + r] = sqrt(top.d[k]);
    }
  }
}

// The flagged sites again, one lane each, in the reference's own candidate
// order (nngp_knn_replay.hpp): MODE 0 the code-book walk, MODE 1 the scan.
// Each lane's list sits in LDS, lane-interleaved (entry k at k * 64 + lane).
constexpr int kReplayBlock = 64;

template <int M, int MODE>
__global__ __launch_bounds__(kReplayBlock) void k_knn_replay(
    KnnReplayArgs A) {
  __shared__ double ld[M * kReplayBlock];
  __shared__ int li[M * kReplayBlock];
  const int t = blockIdx.x * kReplayBlock + (int)threadIdx.x;
  if (t >= A.nflag) return;
  const int site = A.list[t];
  int off, cnt;
  nn_layout(site, A.m, off, cnt);
  double *d = ld + threadIdx.x;
  int *id = li + threadIdx.x;
  if (MODE == 0)
    cb_walk(cnt, A.n, A.ux, A.uy, A.usid, A.upos[site], d, id, kReplayBlock);
  else
    brute_scan(site, cnt, A.x, A.y, d, id, kReplayBlock);
  for (int k = 0; k < cnt; k++) {
    A.nnIndx[off + k] = id[k * kReplayBlock];
    if (A.nnDist) {
      const double v = d[k * kReplayBlock];
      A.nnDist[off + k] = MODE == 0 ? sqrt(v) : v;
    }
  }
}

__device__ __forceinline__ int locate_strip(const StripIndex &S, double x) {
  int lo = 0, hi = S.nstrip - 1;  // last strip with xlo <= x
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S.xlo[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int M>
__global__ __launch_bounds__(kKnnBlock) void k_nn2(Nn2Args A) {
  const int i = blockIdx.x * kKnnBlock + (int)threadIdx.x;
  if (i >= A.nq) return;
  const double2 q = A.q[i];
  TopM<M> top;
  top.init(A.k);
  search<M, true>(A.s, locate_strip(A.s, q.x), q.x, q.y, -1, false, top);
#pragma unroll
  for (int k = 0; k < M; k++) {
    if (k >= M - A.k) {
      const size_t o = (size_t)(k - (M - A.k)) * A.nq + i;
      A.idx[o] = top.id[k];
      if (A.dist) A.dist[o] = sqrt(top.d[k]);
    }
  }
}

int knn_launch(const KnnArgs &a, hipStream_t st) {
  const dim3 grid((a.n + kKnnBlock - 1) / kKnnBlock), block(kKnnBlock);
  if (a.m < 16)
    hipLaunchKernelGGL(k_knn_pred<16>, grid, block, 0, st, a);
  else if (a.m < 33)
    hipLaunchKernelGGL(k_knn_pred<33>, grid, block, 0, st, a);
  else if (a.m < 65)
    hipLaunchKernelGGL(k_knn_pred<65>, grid, block, 0, st, a);
  else
    return -1;
  return (int)hipGetLastError();
}

int knn_replay_launch(const KnnReplayArgs &a, hipStream_t st) {
  if (a.nflag <= 0) return 0;
  const dim3 grid((a.nflag + kReplayBlock - 1) / kReplayBlock),
      block(kReplayBlock);
  if (a.m <= 16) {
    if (a.mode == 0)
      hipLaunchKernelGGL((k_knn_replay<16, 0>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_knn_replay<16, 1>), grid, block, 0, st, a);
  } else if (a.m <= 32) {
    if (a.mode == 0)
      hipLaunchKernelGGL((k_knn_replay<32, 0>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_knn_replay<32, 1>), grid, block, 0, st, a);
  } else if (a.m <= 64) {
    if (a.mode == 0)
      hipLaunchKernelGGL((k_knn_replay<64, 0>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_knn_replay<64, 1>), grid, block, 0, st, a);
  } else {
    return -1;
  }
  return (int)hipGetLastError();
}

// Compacts the sites with flag[i] != 0 into list[0 .. *count), ascending.
int knn_flagged(const unsigned char *flag, int n, int *list, int *count,
                DevBuf &tmp, hipStream_t st) {
  hipcub::CountingInputIterator<int> ids(0);
  size_t bytes = 0;
  NNGP_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, ids, flag, list,
                                         count, n, st));
  NNGP_TRY(tmp.alloc(bytes));
  NNGP_HIP(hipcub::DeviceSelect::Flagged(tmp.p, bytes, ids, flag, list, count,
                                         n, st));
  return NNGP_OK;
}

int knn_build(const double *coords, int n, int m, int mode,
              const HostStrips &h, const DevStrips &ds, const double2 *xy,
              int *nnIndx, double *nnDist, int *lu, hipStream_t st,
              int *nflag_out) {
  DevBuf dFlag, dList, dCount, tmp;
  NNGP_TRY(dFlag.alloc((size_t)n));
  NNGP_TRY(dList.alloc(sizeof(int) * (size_t)n));
  NNGP_TRY(dCount.alloc(sizeof(int)));
  double umax = 0.0;
  for (int i = 0; i < n; i++)
    umax = std::max(umax, std::fabs(coords[i]) + std::fabs(coords[n + i]));
  KnnArgs a;
  a.s = ds.idx;
  a.xy = xy;
  a.strip_of = ds.strip_of.as<int>();
  a.n = n;
  a.m = m;
  a.sorted = h.sorted ? 1 : 0;
  a.nnIndx = nnIndx;
  a.nnDist = nnDist;
  a.lu = lu;
  a.mode = mode;
  a.E = std::ldexp(umax, -48);
  a.flag = dFlag.as<unsigned char>();
  if (knn_launch(a, st)) return set_err(NNGP_ERR_HIP, "kNN launch failed");
  NNGP_TRY(knn_flagged(a.flag, n, dList.as<int>(), dCount.as<int>(), tmp, st));
  int nflag = 0;
  NNGP_HIP(hipMemcpyAsync(&nflag, dCount.p, sizeof(int), hipMemcpyDeviceToHost,
                          st));
  NNGP_HIP(hipStreamSynchronize(st));
  if (nflag_out) *nflag_out = nflag;
  if (nflag == 0) return NNGP_OK;
  KnnReplayArgs r{};
  r.n = n;
  r.m = m;
  r.mode = mode;
  r.nflag = nflag;
  r.list = dList.as<int>();
  r.nnIndx = nnIndx;
  r.nnDist = nnDist;
  DevBuf dux, duy, dus, dup, dc;
  std::vector<double> ux, uy;
  std::vector<int> us, up;
  if (mode == 0) {
    // R's u = x + y order: the permutation rsort_with_index gives
    // (src/nn.cpp:238-243), ties in u included
    std::vector<double> u(n);
    us.resize(n);
    for (int i = 0; i < n; i++) {
      us[i] = i;
      u[i] = coords[i] + coords[n + i];
    }
    r_rsort_with_index(u.data(), us.data(), n, 1);
    ux.resize(n);
    uy.resize(n);
    up.resize(n);
    for (int k = 0; k < n; k++) {
      ux[k] = coords[us[k]];
      uy[k] = coords[n + us[k]];
      up[us[k]] = k;
    }
    NNGP_TRY(dux.alloc(sizeof(double) * (size_t)n));
    NNGP_TRY(duy.alloc(sizeof(double) * (size_t)n));
    NNGP_TRY(dus.alloc(sizeof(int) * (size_t)n));
    NNGP_TRY(dup.alloc(sizeof(int) * (size_t)n));
    NNGP_HIP(hipMemcpyAsync(dux.p, ux.data(), dux.bytes, hipMemcpyHostToDevice,
                            st));
    NNGP_HIP(hipMemcpyAsync(duy.p, uy.data(), duy.bytes, hipMemcpyHostToDevice,
                            st));
    NNGP_HIP(hipMemcpyAsync(dus.p, us.data(), dus.bytes, hipMemcpyHostToDevice,
                            st));
    NNGP_HIP(hipMemcpyAsync(dup.p, up.data(), dup.bytes, hipMemcpyHostToDevice,
                            st));
    r.ux = dux.as<double>();
    r.uy = duy.as<double>();
    r.usid = dus.as<int>();
    r.upos = dup.as<int>();
  } else {
    NNGP_TRY(dc.alloc(sizeof(double) * 2 * (size_t)n));
    NNGP_HIP(hipMemcpyAsync(dc.p, coords, dc.bytes, hipMemcpyHostToDevice, st));
    r.x = dc.as<double>();
    r.y = dc.as<double>() + n;
  }
  if (knn_replay_launch(r, st))
    return set_err(NNGP_ERR_HIP, "kNN replay launch failed");
  // the staging vectors and buffers above must outlive the stream's work
  NNGP_HIP(hipStreamSynchronize(st));
  return NNGP_OK;
}

int nn2_launch(const Nn2Args &a, hipStream_t st) {
  if (a.nq <= 0) return 0;
  const dim3 grid((a.nq + kKnnBlock - 1) / kKnnBlock), block(kKnnBlock);
  if (a.k <= 16)
    hipLaunchKernelGGL(k_nn2<16>, grid, block, 0, st, a);
  else if (a.k <= 32)
    hipLaunchKernelGGL(k_nn2<32>, grid, block, 0, st, a);
  else if (a.k <= 64)
    hipLaunchKernelGGL(k_nn2<64>, grid, block, 0, st, a);
  else
    return -1;
  return (int)hipGetLastError();
}

}  // namespace nngp
