// nngp_strips.hpp — host side of the strip index used by the neighbor
// searches (mkNNIndx(CB) and nn2): points sorted by x, cut into strips of S
// consecutive points, each strip sorted by y.
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "nngp_host.hpp"

namespace nngp {

struct HostStrips {
  int n = 0, S = 0, nstrip = 0;
  std::vector<double> sx, sy, xlo, xhi;
  std::vector<int> sid, strip_of;
  bool sorted = false;
};

// Strip length: the power of two >= sqrt(n m), so a query scans O(sqrt(n m))
// candidates in its own strip and about as many in each neighbor strip.
inline int strip_size(int n, int m) {
  double s = std::sqrt((double)n * (double)std::max(m, 1));
  int S = 64;
  while (S < s && S < (1 << 20)) S <<= 1;
  return std::min(S, std::max(n, 1));
}

// x ascending (ties by id), strips of S, y ascending inside (ties by id)
inline void build_strips(int n, const double *x, const double *y, int m,
                         HostStrips &h) {
  h.n = n;
  h.S = strip_size(n, m);
  h.nstrip = (n + h.S - 1) / h.S;
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  h.sorted = true;
  for (int i = 1; i < n; i++)
    if (x[i] < x[i - 1]) {
      h.sorted = false;
      break;
    }
  if (!h.sorted)
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int b) { return x[a] < x[b]; });
  h.sx.resize(n);
  h.sy.resize(n);
  h.sid.resize(n);
  h.strip_of.resize(n);
  h.xlo.resize(h.nstrip);
  h.xhi.resize(h.nstrip);
  for (int t = 0; t < h.nstrip; t++) {
    const int b = t * h.S, e = std::min(n, b + h.S);
    h.xlo[t] = x[ord[b]];
    h.xhi[t] = x[ord[e - 1]];
    std::vector<int> seg(ord.begin() + b, ord.begin() + e);
    std::sort(seg.begin(), seg.end(), [&](int a, int c) {
      return y[a] < y[c] || (y[a] == y[c] && a < c);
    });
    for (int k = 0; k < e - b; k++) {
      const int id = seg[k];
      h.sx[b + k] = x[id];
      h.sy[b + k] = y[id];
      h.sid[b + k] = id;
      h.strip_of[id] = t;
    }
  }
}

struct DevStrips {
  DevBuf sx, sy, sid, xlo, xhi, strip_of;
  StripIndex idx;
  int upload(const HostStrips &h, hipStream_t st) {
    NNGP_TRY(sx.alloc(sizeof(double) * h.sx.size()));
    NNGP_TRY(sy.alloc(sizeof(double) * h.sy.size()));
    NNGP_TRY(sid.alloc(sizeof(int) * h.sid.size()));
    NNGP_TRY(xlo.alloc(sizeof(double) * h.xlo.size()));
    NNGP_TRY(xhi.alloc(sizeof(double) * h.xhi.size()));
    NNGP_TRY(strip_of.alloc(sizeof(int) * h.strip_of.size()));
    NNGP_HIP(hipMemcpyAsync(sx.p, h.sx.data(), sx.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(sy.p, h.sy.data(), sy.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(sid.p, h.sid.data(), sid.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(xlo.p, h.xlo.data(), xlo.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(xhi.p, h.xhi.data(), xhi.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(strip_of.p, h.strip_of.data(), strip_of.bytes,
                            hipMemcpyHostToDevice, st));
    // the host vectors may be freed by the caller once this returns
    NNGP_HIP(hipStreamSynchronize(st));
    idx.n = h.n;
    idx.S = h.S;
    idx.nstrip = h.nstrip;
    idx.sx = sx.as<double>();
    idx.sy = sy.as<double>();
    idx.sid = sid.as<int>();
    idx.xlo = xlo.as<double>();
    idx.xhi = xhi.as<double>();
    return NNGP_OK;
  }
};

// Compacts the sites with flag[i] != 0 into list[0 .. *count) (device),
// ascending; tmp receives the scan's workspace.
int knn_flagged(const unsigned char *flag, int n, int *list, int *count,
                DevBuf &tmp, hipStream_t st);

// Neighbor lists of n sites (coords: column-major host copy; xy: the same
// on the device, interleaved) into device buffers in the reference layout:
// mode 0 the code-book search (src/nn.cpp:214-257), 1 the scan
// (src/nn.cpp:63-111).  The strip search flags the sites whose tied
// distances need the reference's own candidate order, and those are redone
// in that order.  nnDist may be nullptr; *nflag_out (optional) receives the
// number of redone sites.  Returns after the device work has finished.
int knn_build(const double *coords, int n, int m, int mode,
              const HostStrips &h, const DevStrips &ds, const double2 *xy,
              int *nnIndx, double *nnDist, int *lu, hipStream_t st,
              int *nflag_out);

// Column-major n x 2 coordinates (host) to interleaved double2 (device).
inline int upload_xy(const double *cm, int n, DevBuf &xy, hipStream_t st) {
  DevBuf tmp;
  NNGP_TRY(tmp.alloc(sizeof(double) * 2 * (size_t)n));
  NNGP_HIP(hipMemcpyAsync(tmp.p, cm, tmp.bytes, hipMemcpyHostToDevice, st));
  NNGP_TRY(xy.alloc(sizeof(double2) * (size_t)n));
  if (interleave_launch(tmp.as<double>(), n, xy.as<double2>(), st))
    return set_err(NNGP_ERR_HIP, "coordinate upload kernel failed");
  NNGP_HIP(hipStreamSynchronize(st));
  return NNGP_OK;
}

}  // namespace nngp
