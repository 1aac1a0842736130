// nngp_knn.hip — exact nearest-neighbor search on gfx950.
//
// mkNNIndxCB (src/nn.cpp:128-257) finds, for every site i, the m nearest
// sites among its NNGP predecessors (index < i), ascending by the squared
// distance fl(fl(dx*dx) + fl(dy*dy)), then stores sqrt of it.  The reference
// walks a u = x + y sorted list; this search uses a strip index (points
// sorted by x, cut into strips, each strip sorted by y) and one query per
// lane.  Every pruning test compares a rounded partial square (dx^2 of the
// strip gap, or dy^2 inside a strip) with the current m-th distance; since
// fl(dx^2) + fl(dy^2) rounds monotonically, nothing that could enter the
// list is ever skipped, so the result is the exact m smallest set in the
// reference's own arithmetic.  Products and sums here must not contract
// into FMAs (the reference rounds dx*dx, dy*dy and their sum separately), so
// contraction is switched off for this whole file.
//
// Insertion is strict (d2 < current m-th), an equal distance goes after the
// ones already held.  Exactly tied distances may be ordered differently from
// the reference (whose order depends on its walk and its shell sort).
#include "nngp_kernels.hpp"

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

// Scan strip t; `self` >= 0 restricts candidates to ids < self.
template <int M>
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
  // false once dy^2 alone exceeds the current m-th distance
  auto visit = [&](int j) {
    const double dy = qy - sy[j];
    const double dy2 = dy * dy;
    if (dy2 > top.worst()) return false;
    const int id = sid[j];
    if (self < 0 || id < self) {
      const double dx = qx - sx[j];
      const double dx2 = dx * dx;
      const double d2 = dx2 + dy2;
      if (d2 < top.worst()) top.insert(d2, id);
    }
    return true;
  };
  for (int j = p; j < len; j++)
    if (!visit(j)) break;
  for (int j = p - 1; j >= 0; j--)
    if (!visit(j)) break;
}

template <int M>
__device__ inline void search(const StripIndex &S, int t0, double qx,
                              double qy, int self, bool left_only,
                              TopM<M> &top) {
  scan_strip<M>(S, t0, qx, qy, self, top);
  for (int t = t0 - 1; t >= 0; t--) {
    const double gap = qx - S.xhi[t];
    if (gap > 0.0 && gap * gap > top.worst()) break;
    scan_strip<M>(S, t, qx, qy, self, top);
  }
  if (left_only) return;
  for (int t = t0 + 1; t < S.nstrip; t++) {
    const double gap = S.xlo[t] - qx;
    if (gap > 0.0 && gap * gap > top.worst()) break;
    scan_strip<M>(S, t, qx, qy, self, top);
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

template <int M>
__global__ __launch_bounds__(kKnnBlock) void k_knn_pred(KnnArgs A) {
  const int i = blockIdx.x * kKnnBlock + (int)threadIdx.x;
  if (i >= A.n) return;
  int off, cnt;
  nn_layout(i, A.m, off, cnt);
  A.lu[i] = off;
  A.lu[A.n + i] = cnt;
  if (cnt == 0) return;
  const double2 q = A.xy[i];
  TopM<M> top;
  top.init(cnt);
  search<M>(A.s, A.strip_of[i], q.x, q.y, i, A.sorted != 0, top);
#pragma unroll
  for (int k = 0; k < M; k++) {
    if (k >= M - cnt) {
      const int o = off + k - (M - cnt);
      A.nnIndx[o] = top.id[k];
      if (A.nnDist) A.nnDist[o] = sqrt(top.d[k]);
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
  search<M>(A.s, locate_strip(A.s, q.x), q.x, q.y, -1, false, top);
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
  if (a.m <= 16)
    hipLaunchKernelGGL(k_knn_pred<16>, grid, block, 0, st, a);
  else if (a.m <= 32)
    hipLaunchKernelGGL(k_knn_pred<32>, grid, block, 0, st, a);
  else if (a.m <= 64)
    hipLaunchKernelGGL(k_knn_pred<64>, grid, block, 0, st, a);
  else
    return -1;
  return (int)hipGetLastError();
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
