// nngp_kernels.hpp — launchers of the NNGP device kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/nngp.h"
#include "nngp_device.hpp"

namespace nngp {

constexpr int kMaxP = 8;    // covariates supported by the fused passes
constexpr int kMaxNZ = kMaxP + 1;
constexpr int kStatsMax = 2 + kMaxNZ * (kMaxNZ + 1) / 2;  // logdet, Q, gram

// One neighbor-system pass over sites [0, n): B, F of one theta, and the
// per-site terms summed per block: column 0 sum_i log F_i, column 1 zero,
// then the lower triangle (row by row) of sum_i e_i e_i^T / F_i where e_i is
// the residual vector (X row, y) - B_i (X, y)_N(i).  Quadratic forms of any
// r = X beta - y follow from these cross-products.
struct BfArgs {
  const double2 *xy;     // n interleaved coordinates
  const int *nnIndx;     // flat neighbor index (+64 readable padding)
  const int *lu;         // nnIndxLU, 2n
  int n;
  int m;                 // n.neighbors (max count of any site)
  double *B;             // nIndx, or nullptr: not stored
  double *F;             // n, or nullptr: not stored
  const double *Z;       // n x zs row-major: X row then y (nz = p + 1 used)
  int nz;
  int zs;
  int *err;              // smallest site whose Cholesky failed (INT_MAX = ok)
  double *rows;          // partial rows, column k at rows[k*rows_cap + r],
  int rows_cap;          // or nullptr when no sums are wanted
  CovArgs c;
  int variant;           // cov_variant(c.cov, c.nu): templated kernel, or -1
  int mixed = 0;         // 1: FP32 factor/solve + one FP64 refinement step
                         // (per-lane shape, single pass only; else FP64)
  // Batched passes (ng > 0): gridDim.y problems, problem j = fold j / ng at
  // grid point j % ng, whose covariance parameters are covs[j % ng]; fold f
  // reads xy + f*fs_xy, nnIndx + f*fs_nn, lu + f*fs_lu, Z + f*fs_Z; problem
  // j writes rows + j*ps_rows and err[j].  B and F are not stored.
  int ng = 0;
  const CovArgs *covs = nullptr;
  long long fs_xy = 0, fs_nn = 0, fs_lu = 0, fs_Z = 0, ps_rows = 0;
  int coop_blocks = 0;   // set by bf_launch
  // 1: lu follows the closed form of src/util.cpp:55-70 (offset
  // i(i-1)/2 for i < m, else m(m-1)/2 + (i-m)m; count min(i, m)), checked
  // on the host, so the per-lane kernel computes it instead of loading it
  int lu_closed = 0;
};

// Offset and count of site i in the flat neighbor arrays when they follow
// the closed form (src/util.cpp:55-70).
__host__ __device__ inline void nn_closed(int i, int m, int &off, int &cnt) {
  if (i < m) {
    cnt = i;
    off = i * (i - 1) / 2;
  } else {
    cnt = m;
    off = m * (m - 1) / 2 + (i - m) * m;
  }
}

// Launch shape of a pass: family 0 = one site per lane (threads 64, 128 or
// 256 per block), family 1 = one site per wavefront (threads = 64 x waves,
// waves 1, 2 or 4, each wave an LDS tile of one system), family 2 = one
// site per 16-lane row, four per wavefront (m <= 32, single FP64 problems;
// threads 64, 128 or 256).
struct BfConfig {
  int family = 0;
  int threads = 256;
  int mode = 0;          // mixed-precision passes: 1 = covariance copy in
                         // LDS, 3 = in registers; FP64 per-lane passes:
                         // 2 = covariate rows staged in LDS
};

struct BfTuneReport {
  int n = 0;             // candidates timed (at most NNGP_TUNE_MAX kept)
  int family[NNGP_TUNE_MAX], threads[NNGP_TUNE_MAX], mode[NNGP_TUNE_MAX];
  double ms[NNGP_TUNE_MAX];  // mean launch time of each
  BfConfig best;
  double best_ms = 0.0;
};

// Plain Q(u, v) of src/util.cpp:261-280 over stored B and F.
struct QArgs {
  const double *B, *F, *u, *v;
  const int *nnIndx, *lu;
  int n;
  double *rows;          // kRedBlocks partial rows
  double *out;
};

// Partial sums k_quadform writes: one per block of its grid-stride launch,
// never more than this many, whatever n.
constexpr int kQPartialRows = 1024;
int q_partial_rows(int n);

// Rows a partial-row buffer needs for n sites under any launch shape.
int rows_capacity(int n);
bool bf_lane_ok(int m, int variant);
bool bf_config_ok(const BfConfig &cfg, int m, int variant);
// Untuned launch shape: the per-lane kernel where it applies, else the row
// kernel for single FP64 problems (single_fp64) with m <= 32, else the
// wavefront kernel's widest LDS tile.
BfConfig bf_default(int m, int variant, bool single_fp64);
// Partial rows one problem of a pass writes under cfg.
int bf_rows(const BfArgs &a, const BfConfig &cfg);
// nprob problems (1 unless a.ng > 0).  -1: cfg does not apply to (m,
// variant); -2: bad batch or rows_cap too small; else a hipError_t.
int bf_launch(const BfArgs &a, const BfConfig &cfg, int nprob,
              hipStream_t st);
// Completion word of a sum launch in host memory: the launch's last block
// to finish writes seq to *flag (a device address of host memory) once the
// outputs of every block are visible to the host, so the host can poll it
// instead of synchronizing the stream.  done: a device counter, zero
// between launches (the last block resets it).
struct SumsSignal {
  int *done = nullptr;
  unsigned *flag = nullptr;
  unsigned seq = 0;
};
// Sums of the bf_launch with the same arguments: problem j's column k to
// out[j*out_stride + k]; its failure flag to err_out[j] (and reset).
int bf_sums_launch(const BfArgs &a, const BfConfig &cfg, int nprob,
                   double *out, int out_stride, double *err_out,
                   hipStream_t st, const SumsSignal *sig = nullptr);
// Column sums of nprob blocks of partial rows (see k_rows_sum).
int rows_sum_launch(const double *rows, int cap, int nrows, long long ps_rows,
                    int nv, int nprob, double *out, int out_stride, int *err,
                    double *err_out, hipStream_t st,
                    const SumsSignal *sig = nullptr);
// Time every applicable launch shape of this pass and keep the fastest
// (cached per (m, variant, nz, size class) unless rep asks for timings).
int bf_autotune(const BfArgs &a, int nprob, hipStream_t st, BfConfig *best,
                BfTuneReport *rep);
int q_launch(const QArgs &a, hipStream_t st);

// *p = v on the stream (no memset node, safe under graph capture).
int fill_int_launch(int *p, int v, hipStream_t st);

// Interleave column-major coords (x block then y block) into xy.
int interleave_launch(const double *coords_cm, int n, double2 *xy,
                      hipStream_t st);

// ---------------------------------------------------------------------------
// neighbor search
// ---------------------------------------------------------------------------
// Strip index: points sorted by x, cut into strips of S consecutive points,
// each strip sorted by y.  Arrays are strip-major (strip t at [t*S, ...)).
struct StripIndex {
  int n, S, nstrip;
  const double *sx, *sy;  // coordinates
  const int *sid;         // point id (NNGP index, or fitted row for nn2)
  const double *xlo, *xhi;  // x range of each strip
};

// mkNNIndx(CB): m nearest predecessors of every site i (ids < i).
struct KnnArgs {
  StripIndex s;
  const double2 *xy;     // NNGP-order coordinates
  const int *strip_of;   // strip holding each NNGP index
  int n, m;
  int sorted;            // 1: NNGP order equals the x order
  int *nnIndx;           // outputs in the reference layout
  double *nnDist;        // or nullptr: not stored
  int *lu;
  int mode;              // 0: code-book walk (mkNNIndxCB), 1: scan (mkNNIndx)
  double E;              // 2^-48 max(|x| + |y|): rounding scale of the walk
  unsigned char *flag;   // n: 1 where ties need the reference's own order
};
int knn_launch(const KnnArgs &a, hipStream_t st);

// The flagged sites redone in the reference's candidate order.
struct KnnReplayArgs {
  int n, m, mode, nflag;
  const int *list;                 // flagged site ids
  const double *ux, *uy;           // mode 0: coordinates in R's u order
  const int *usid, *upos;          //   site at each position, and inverse
  const double *x, *y;             // mode 1: coordinates by site
  int *nnIndx;
  double *nnDist;
};
int knn_replay_launch(const KnnReplayArgs &a, hipStream_t st);

// nn2: k nearest fitted points of each query point.
struct Nn2Args {
  StripIndex s;
  const double2 *q;      // query coordinates
  int nq, k;
  int *idx;              // nq x k column-major (0-based)
  double *dist;          // nq x k column-major or nullptr
};
int nn2_launch(const Nn2Args &a, hipStream_t st);

// ---------------------------------------------------------------------------
// kriging
// ---------------------------------------------------------------------------
// One run of consecutive sites of one kriging request inside a batched launch
// (the serving loop): blocks [block0, block0 + count) of the launch handle
// the request's sites first .. first + count - 1.
struct PredSeg {
  const double2 *xy0;    // the request's q coordinates
  const double *X0;      // q x p column-major
  const int *nn0;        // q x m column-major
  double *y0;            // q x S: y0[s*q + i]
  int *err;              // the request's failure flag
  unsigned long long seed;  // z(i, s) = philox_normal(seed, i, s)
  int q, first, block0, count;
};

struct PredArgs {
  const double2 *xy;     // fitted coordinates (original order)
  const double *X;       // n x p column-major
  const double *y;       // n
  int n, p, m;
  const double2 *xy0;    // q query coordinates
  const double *X0;      // q x p column-major
  int q;
  const int *nn0;        // q x m column-major
  const double *beta;    // p x S
  const double *theta;   // nTheta x S
  int S, nTheta, cov;
  const double *z;       // q*S normals, z[i*S + s]
  double *y0;            // q x S: y0[s*q + i]
  int *err;
  int site0;             // first site of this launch (sites [site0, site0+nsite))
  int nsite;
  CovArgs unit;          // scale-free exp coefficients (closed-form path)
  const CovArgs *cov_s;  // per-sample parameters, device, S entries
                         // (generic path)
  // Batched form: when segs != nullptr the launch covers nsite blocks split
  // into nseg runs (sorted by block0); xy0, X0, nn0, q, z, y0, err and site0
  // above are then unused and z comes from each run's seed.
  const PredSeg *segs = nullptr;
  int nseg = 0;
  // Latent form (src/sNNGPPredict.cpp:163-205) when wS != nullptr: C has no
  // nugget, the weights apply to the posterior w samples wS (n x S,
  // column s = sample s), w0 = d + sqrt(s2 - w.c) z and
  // y0 = sqrt(t2) z2 + x0.beta + w0 with a second set of q*S normals z2
  // (z2[i*S + s]); w0 goes to w0out (q x S).  Unbatched launches only.
  const double *wS = nullptr;
  const double *z2 = nullptr;
  double *w0out = nullptr;
};
int pred_launch(const PredArgs &a, hipStream_t st);

struct ConjPredArgs {
  const double2 *xy;
  const double *X, *y;
  int n, p, m;
  const double2 *xy0;
  const double *X0;
  int n0;
  const int *nn0;        // n0 x m column-major
  const double *beta;    // p
  const double *Vb;      // p x p (lower) = (X^T V^-1 X)^-1
  double a, b, alpha;
  double *y0, *y0var;    // n0
  int *err;
  CovArgs c;
};
int conj_pred_launch(const ConjPredArgs &a, hipStream_t st);

}  // namespace nngp
