// nngp_host.hpp — host-side helpers of libnngp_hip.so: error state, device
// buffers, covariance parameters, small dense algebra and the built-in
// random number generator.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nngp.h"
#include "nngp_kernels.hpp"

namespace nngp {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
int set_err(int code, const char *fmt, ...);
void clear_err();

#define NNGP_HIP(expr)                                                      \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ::nngp::set_err(NNGP_ERR_HIP, "HIP error '%s' (%s:%d)",        \
                             hipGetErrorString(e_), __FILE__, __LINE__);    \
  } while (0)

#define NNGP_TRY(expr)            \
  do {                            \
    int st_ = (expr);             \
    if (st_ != NNGP_OK) return st_; \
  } while (0)

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t nbytes) {
    release();
    if (nbytes == 0) nbytes = 8;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
      p = nullptr;
      return set_err(e == hipErrorOutOfMemory ? NNGP_ERR_NOMEM : NNGP_ERR_HIP,
                     "device allocation of %zu bytes failed: %s", nbytes,
                     hipGetErrorString(e));
    }
    bytes = nbytes;
    return NNGP_OK;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct PinnedBuf {
  void *p = nullptr;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  int alloc(size_t nbytes) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    NNGP_HIP(hipHostMalloc(&p, nbytes ? nbytes : 8, hipHostMallocDefault));
    return NNGP_OK;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() = default;
  Stream(const Stream &) = delete;
  Stream &operator=(const Stream &) = delete;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
  int create() {
    NNGP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return NNGP_OK;
  }
};

// Select the device of opts (default 0); fails cleanly without a GPU.
int use_device(const nngp_opts *opts);

// ---------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------
CovArgs make_cov(int cov, double s2, double t2, double phi, double nu);

// ---------------------------------------------------------------------------
// argument checks
// ---------------------------------------------------------------------------
long long n_indx(int n, int m);
// nnIndxLU consistent with nIndx, counts <= m, indices in [0, n).
int check_neighbors(int n, int m, const int *nnIndx, const int *nnIndxLU);

// ---------------------------------------------------------------------------
// small dense algebra (lower triangle, column-major, the reference's
// dpotrf / dpotri / dsymv / dtrmv sequence of src/rNNGP.cpp:288-294): LAPACK
// dpotf2 / dtrti2 + dlauu2, BLAS dsymv / dtrmv loop orders, lda = n; host
// and device (the resident chain), without contracted multiply-adds
// ---------------------------------------------------------------------------
#define A_(i, j) A[(size_t)(j) * n + (i)]
__host__ __device__ inline int h_potrf_L(int n, double *A) {
#pragma clang fp contract(off)
  for (int j = 0; j < n; j++) {
    double dot = 0.0;
    for (int k = 0; k < j; k++) dot += A_(j, k) * A_(j, k);
    double ajj = A_(j, j) - dot;
    if (ajj <= 0.0 || ajj != ajj) return j + 1;
    ajj = sqrt(ajj);
    A_(j, j) = ajj;
    for (int k = 0; k < j; k++) {
      const double t = -A_(j, k);
      for (int i = j + 1; i < n; i++) A_(i, j) += t * A_(i, k);
    }
    const double r = 1.0 / ajj;
    for (int i = j + 1; i < n; i++) A_(i, j) *= r;
  }
  return 0;
}

__host__ __device__ inline int h_potri_L(int n, double *A) {
#pragma clang fp contract(off)
  for (int j = 0; j < n; j++)
    if (A_(j, j) == 0.0) return j + 1;
  for (int j = n - 1; j >= 0; j--) {
    A_(j, j) = 1.0 / A_(j, j);
    const double ajj = -A_(j, j);
    const int nn = n - j - 1;
    for (int jj = nn - 1; jj >= 0; jj--) {
      const double t = A_(j + 1 + jj, j);
      if (t != 0.0) {
        for (int ii = nn - 1; ii > jj; ii--)
          A_(j + 1 + ii, j) += t * A_(j + 1 + ii, j + 1 + jj);
        A_(j + 1 + jj, j) = t * A_(j + 1 + jj, j + 1 + jj);
      }
    }
    for (int ii = 0; ii < nn; ii++) A_(j + 1 + ii, j) *= ajj;
  }
  for (int i = 0; i < n; i++) {
    const double aii = A_(i, i);
    if (i < n - 1) {
      double dot = 0.0;
      for (int k = i; k < n; k++) dot += A_(k, i) * A_(k, i);
      A_(i, i) = dot;
      for (int jj = 0; jj < i; jj++) {
        double t = 0.0;
        for (int k = i + 1; k < n; k++) t += A_(k, jj) * A_(k, i);
        A_(i, jj) = t + aii * A_(i, jj);
      }
    } else {
      for (int jj = 0; jj <= i; jj++) A_(i, jj) *= aii;
    }
  }
  return 0;
}

__host__ __device__ inline void h_symv_L(int n, const double *A,
                                         const double *x, double *y) {
#pragma clang fp contract(off)
  for (int i = 0; i < n; i++) y[i] = 0.0;
  for (int j = 0; j < n; j++) {
    const double t1 = x[j];
    double t2 = 0.0;
    y[j] += t1 * A_(j, j);
    for (int i = j + 1; i < n; i++) {
      y[i] += t1 * A_(i, j);
      t2 += A_(i, j) * x[i];
    }
    y[j] += t2;
  }
}

__host__ __device__ inline void h_trmv_L(int n, const double *A, double *x) {
#pragma clang fp contract(off)
  for (int j = n - 1; j >= 0; j--) {
    const double t = x[j];
    if (t != 0.0) {
      for (int i = n - 1; i > j; i--) x[i] += t * A_(i, j);
      x[j] = t * A_(j, j);
    }
  }
}
#undef A_

// ---------------------------------------------------------------------------
// random numbers with R's argument semantics
// ---------------------------------------------------------------------------
struct BuiltinRng {
  // xoshiro256** (Blackman & Vigna) seeded through splitmix64; normals by
  // the polar Box-Muller method, one cached.
  uint64_t s[4];
  int have = 0;
  double cached = 0.0;
  explicit BuiltinRng(uint64_t seed);
  uint64_t next();
  double unif();
  double norm();
};

struct Rng {
  const nngp_rng *ext = nullptr;
  BuiltinRng own;
  explicit Rng(const nngp_rng *r, uint64_t seed) : ext(r), own(seed) {}
  double norm_rand() { return ext ? ext->norm_rand(ext->state) : own.norm(); }
  double unif_rand() { return ext ? ext->unif_rand(ext->state) : own.unif(); }
  // R's rnorm(mu, sd): sd == 0 or non-finite mu returns mu without a draw
  double rnorm(double mu, double sd) {
    if (std::isnan(mu) || !std::isfinite(sd) || sd < 0.0) return NAN;
    if (sd == 0.0 || !std::isfinite(mu)) return mu;
    return mu + sd * norm_rand();
  }
  // R's runif(a, b)
  double runif(double a, double b) {
    if (!std::isfinite(a) || !std::isfinite(b) || b < a) return NAN;
    if (a == b) return a;
    double u;
    do {
      u = unif_rand();
    } while (u <= 0.0 || u >= 1.0);
    return a + (b - a) * u;
  }
};

// ---------------------------------------------------------------------------
// device-resident model
// ---------------------------------------------------------------------------
// device sums buffer: two passes of kStatsMax values, then their two
// updateBF failure flags
constexpr int kSumsLen = 2 * kStatsMax + 2;

struct PassOut {
  double logDet = 0.0, q = 0.0;
  double gram[kStatsMax];  // the raw sums: [0] logDet, then from [2] the
                           // lower triangle of E^T F^-1 E, row-packed
};

// X^T V^-1 X entry (a, b), X^T V^-1 y (a, p) or y^T V^-1 y (p, p) of a pass
double gram_at(const double *raw, int a, int b);
// Q(r, r) for r = X beta - y: beta' G beta - 2 beta' h + y' V^-1 y
double quad_resid(const PassOut &po, const double *beta, int p);

// One replicate draw v_i = B_i . v_N(i) + sqrt(F_i) z_i (i = 0 .. n-1, the
// updateRep of src/rNNGP.cpp:24-42) on the device from device B and F,
// with the caller's n normals: the caller fills z() (pinned) in site order,
// draw() uploads them, walks the neighbor DAG level by level (the level
// schedule built once by init from the host index) and leaves v in v().
struct RepDrawer {
  int n = 0, m = 0, nlev = 0;  // m: the largest neighbor count
  DevBuf dorder, dloff, dsites, dz, dv, dy;
  PinnedBuf hz, hv;
  int init(int n, const int *nnIndx, const int *nnIndxLU);
  double *z() { return hz.as<double>(); }
  const double *v() const { return hv.as<double>(); }
  int draw(const double *B, const double *F, const int *nnIndx,
           const int *lu, hipStream_t st);
};

struct Model {
  int n = 0, m = 0, p = 0, nz = 0, zs = 0, cov = 0, precision = 0;
  long long nIndx = 0;
  Stream st;
  DevBuf xy, nnIndx, lu, Z;
  DevBuf B[2], F[2];
  // sums: [0, kStatsMax) pass at off 0, [kStatsMax, 2 kStatsMax) pass at
  // off kStatsMax, then the updateBF failure flag of each (as doubles)
  DevBuf rows, sums, err;
  PinnedBuf hsums;  // host mirror of sums, then the completion word
  double *sums_mapped = nullptr;  // device address of hsums: the sum kernel
                                  // writes there directly (no copy)
  // With sums_mapped, each sum launch ends by writing its sequence number
  // to the word after the sums (SumsSignal), which fetch_sums polls.
  DevBuf sig_done;        // the sum kernels' block counter
  unsigned sig_seq = 0;   // sequence number of the last sum launch
  BfArgs last_bf;      // arguments of the last bf() launch
  BfConfig last_cfg;   // and its launch shape
  BfConfig tuned_cfg;  // launch shape chosen by autotune(), when have_tuned
  bool have_tuned = false;
  bool lu_closed = false;  // nnIndxLU follows the closed form (nn_closed)
  PassOut slot_out[2]; // sums of the last completed pass into each slot

  // Time the launch shapes of the pass at c and use the fastest from now on.
  int autotune(const CovArgs &c, BfTuneReport *rep);
  // Use this launch shape from now on (NNGP_ERR_ARG if it does not apply).
  int pin_shape(int family, int threads, int mode);

  int create(int n, int m, int p, const double *coords, const int *nnIndx,
             const int *nnIndxLU, const double *X, const double *y, int cov,
             const nngp_opts *opts);
  // B, F and the block sums of theta into slot; B stored only when storeB
  int bf(const CovArgs &c, int slot, bool storeB);
  // sums of the last bf() into sums[off .. off + kStatsMax)
  int sums_launch(int slot, int off);
  // one full pass: bf + sums, synchronized; out (and slot_out[slot])
  int pass(const CovArgs &c, int slot, bool storeB, PassOut *out);
  // copy sums (and the failure flags) to hsums and synchronize
  int fetch_sums();
  // failure flag of the pass whose sums sit at off 0 (which = 0) or at off
  // kStatsMax (which = 1), as last fetched
  int check_err(const char *what, int which);
  void unpack(const double *raw, PassOut *out) const;

  // optional kernel timing with HIP events on st (kind 0: updateBF,
  // kind 1: sum kernels); collect() after each stream synchronization.
  // Every prof_every-th updateBF launch and the sum launch after it are
  // timed: each event is a packet on the stream, and timing every launch
  // slows a chain of short passes by about 5%.
  struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
    int kind = 0;
  };
  bool prof = false;
  int prof_every = 1;
  long long prof_seq = 0;
  bool prof_sample = false;  // the last updateBF launch was timed
  std::vector<EvPair> evs;
  int nev = 0;
  double tms[2] = {0.0, 0.0};
  long long tcount[2] = {0, 0};
  // host wall time of the samplers' iteration loops while timing is on:
  // from the first proposal (initial pass done, stream drained) to the last
  // decision (stream drained); the initial pass is outside
  double loop_ms = 0.0;
  long long loop_iters = 0;
  int ev_begin(int kind);
  void ev_end(int idx);
  int collect();
  ~Model();
};

}  // namespace nngp
