// nngp_knn_replay.hpp — the reference's own neighbor-search procedures,
// replayed step for step for the sites whose result depends on how exactly
// tied distances are visited and sorted.
//
// Both of the reference's searches keep a list of the m best candidates,
// accept a candidate only when its distance is strictly below the list's
// last entry, write it over that entry and re-sort the list with R's
// rsort_with_index.  That sort is a shell sort, so it is not stable: where
// distances tie, the order of the list (and, at the m-th place, its
// membership) depends on the order the candidates arrive in and on every
// intermediate state of the list.  The strip search of nngp_knn.hip finds
// the exact m smallest distances and flags every site where two of its m+1
// smallest are equal or closer than the rounding of the reference's
// pruning test; those sites are redone here, one lane each, in the
// reference's own candidate order:
//   - cb_walk:    the code-book walk of src/nn.cpp:136-211 (candidates in
//                 the u = x + y order, alternating down and up, a direction
//                 dropped once (du)^2 > 2 * worst);
//   - brute_scan: the scan of src/nn.cpp:94-107 (candidates j = 0 .. i-1,
//                 Euclidean distances compared).
// The list lives in memory with a stride (LDS, lane-interleaved, on the
// device).  Each function switches contraction off for its own body: the
// reference rounds each square and sum separately, and a fused multiply-add
// would change which candidates tie.
#pragma once

namespace nngp {

// R's rsort_with_index (R's src/main/sort.c): shell sort over the gaps
// 1, 4, 13, 40, ... starting at the first gap above n / 9; an entry moves
// only past strictly greater keys.  Entry k of the list is x[k * st].
__host__ __device__ inline void r_rsort_with_index(double *x, int *ix, int n,
                                                   int st) {
  int h = 1;
  while (h <= n / 9) h = 3 * h + 1;
  for (; h > 0; h /= 3) {
    for (int i = h; i < n; i++) {
      const double v = x[i * st];
      const int iv = ix[i * st];
      int j = i;
      while (j >= h && x[(j - h) * st] > v) {
        x[j * st] = x[(j - h) * st];
        ix[j * st] = ix[(j - h) * st];
        j -= h;
      }
      x[j * st] = v;
      ix[j * st] = iv;
    }
  }
}

// src/nn.cpp:136-211: the cnt nearest predecessors of the site at position
// ui of the u-sorted order.  ux, uy are the coordinates in that order, usid
// the site ids; on return d[k * st] holds squared distances, ascending.
template <class Real>
__host__ __device__ inline void cb_walk(int cnt, int n, const Real *ux,
                                        const Real *uy, const int *usid,
                                        int ui, double *d, int *id, int st) {
#pragma clang fp contract(off)
  for (int k = 0; k < cnt; k++) {
    d[k * st] = __builtin_inf();
    id[k * st] = 0;
  }
  if (cnt == 0) return;
  const double xq = ux[ui], yq = uy[ui];
  const int q = usid[ui];
  const int last = (cnt - 1) * st;
  double worst = __builtin_inf();
  int i = ui, j = ui;
  bool up = true, down = true;
  while (up || down) {
    if (i == 0) down = false;
    if (j == n - 1) up = false;
    if (down) {
      i--;
      const double xc = ux[i], yc = uy[i];
      const double t = xq + yq - xc - yc;  // ((xq + yq) - xc) - yc
      if (t * t > 2 * worst) {
        down = false;
      } else {
        const double a = xq - xc, b = yq - yc;
        const double de = a * a + b * b;
        const int c = usid[i];
        if (de < worst && c < q) {
          d[last] = de;
          id[last] = c;
          r_rsort_with_index(d, id, cnt, st);
          worst = d[last];
        }
      }
    }
    if (up) {
      j++;
      const double xc = ux[j], yc = uy[j];
      const double t = xq + yq - xc - yc;
      if (t * t > 2 * worst) {
        up = false;
      } else {
        const double a = xq - xc, b = yq - yc;
        const double de = a * a + b * b;
        const int c = usid[j];
        if (de < worst && c < q) {
          d[last] = de;
          id[last] = c;
          r_rsort_with_index(d, id, cnt, st);
          worst = d[last];
        }
      }
    }
  }
}

// src/nn.cpp:94-107: the cnt nearest of sites 0 .. i-1 to site i by the
// Euclidean distance sqrt((xi-xj)^2 + (yi-yj)^2) (src/util.cpp:51-53);
// d[k * st] holds the distances, ascending.
template <class Real>
__host__ __device__ inline void brute_scan(int i, int cnt, const Real *x,
                                           const Real *y, double *d, int *id,
                                           int st) {
#pragma clang fp contract(off)
  for (int k = 0; k < cnt; k++) {
    d[k * st] = __builtin_inf();
    id[k * st] = 0;
  }
  if (cnt == 0) return;
  const double xi = x[i], yi = y[i];
  const int last = (cnt - 1) * st;
  double worst = __builtin_inf();
  for (int j = 0; j < i; j++) {
    const double a = xi - x[j], b = yi - y[j];
    const double dist = sqrt(a * a + b * b);
    if (dist < worst) {
      d[last] = dist;
      id[last] = j;
      r_rsort_with_index(d, id, cnt, st);
      worst = d[last];
    }
  }
}

}  // namespace nngp
