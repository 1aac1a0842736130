// nngp_host.hip — host helpers and the device-resident model of
// libnngp_hip.so (see nngp_host.hpp).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>

#include "nngp_host.hpp"

namespace nngp {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int set_err(int code, const char *fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
void clear_err() { g_err.clear(); }
const char *last_error() { return g_err.c_str(); }

int use_device(const nngp_opts *opts) {
  const int dev = opts ? opts->device : 0;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return set_err(NNGP_ERR_HIP, "no HIP device available (%s)",
                   e != hipSuccess ? hipGetErrorString(e) : "0 devices");
  if (dev < 0 || dev >= count)
    return set_err(NNGP_ERR_ARG, "device %d out of range (%d visible)", dev,
                   count);
  NNGP_HIP(hipSetDevice(dev));
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// covariance parameters
// ---------------------------------------------------------------------------

// Coefficients of 1/Gamma(z) = sum_k c_k z^k (Abramowitz & Stegun 6.1.34),
// even k, for Gamma_1(mu) = -(c2 + c4 mu^2 + ...) near mu = 0.
static const double kRg[] = {0.5772156649015329,  -0.0420026350340952,
                             -0.0421977345555443, 0.0072189432466630,
                             -0.0002152416741149, -0.0000201348547807,
                             0.0000011330272320};

// Coefficients of the degree-10 polynomial that interpolates 2^x at the 11
// Chebyshev nodes of [-1/2, 1/2] (Gaussian elimination on the Vandermonde
// system in long double).  Interpolation at Chebyshev nodes is within a
// small factor of the minimax polynomial: relative error 4.5e-16 on the
// interval, against 8.7e-15 for the degree-11 Taylor polynomial of exp(f ln 2).
static const long double *exp2_cheb() {
  static long double c[kExp2Deg + 1];
  static const bool done = [] {
    constexpr int n = kExp2Deg + 1;
    long double A[n][n], b[n];
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; k++) {
      const long double x = 0.5L * cosl((2 * k + 1) * pi / (2 * n));
      long double t = 1.0L;
      for (int j = 0; j < n; j++) {
        A[k][j] = t;
        t *= x;
      }
      b[k] = exp2l(x);
    }
    for (int i = 0; i < n; i++) {
      int pv = i;
      for (int r = i + 1; r < n; r++)
        if (fabsl(A[r][i]) > fabsl(A[pv][i])) pv = r;
      for (int j = 0; j < n; j++) std::swap(A[i][j], A[pv][j]);
      std::swap(b[i], b[pv]);
      for (int r = i + 1; r < n; r++) {
        const long double f = A[r][i] / A[i][i];
        for (int j = i; j < n; j++) A[r][j] -= f * A[i][j];
        b[r] -= f * b[i];
      }
    }
    for (int i = n - 1; i >= 0; i--) {
      long double acc = b[i];
      for (int j = i + 1; j < n; j++) acc -= A[i][j] * c[j];
      c[i] = acc / A[i][i];
    }
    return true;
  }();
  (void)done;
  return c;
}

CovArgs make_cov(int cov, double s2, double t2, double phi, double nu) {
  CovArgs c;
  std::memset(&c, 0, sizeof c);
  c.cov = cov;
  c.s2 = s2;
  c.t2 = t2;
  c.phi = phi;
  c.nu = nu;
  c.diag = s2 * 1.0 + t2;
  c.inv_phi = 1.0 / phi;
  c.phi2 = phi * phi;
  const long double ln2 = logl(2.0L);
  double hi = (double)ln2;
  uint64_t bits;
  std::memcpy(&bits, &hi, 8);
  bits &= 0xFFFFFFFF00000000ull;  // 21 significant bits: k*hi is exact
  std::memcpy(&hi, &bits, 8);
  c.ln2_hi = hi;
  c.ln2_lo = (double)(ln2 - (long double)hi);
  c.log2e = (double)(1.0L / ln2);
  double fact = 1.0;
  const long double *ch = exp2_cheb();
  for (int k = 0; k <= kExpDeg; k++) {
    if (k > 0) fact *= k;
    c.ec[k] = s2 / fact;
    c.ec2[k] = k <= kExp2Deg ? (double)((long double)s2 * ch[k]) : 0.0;
  }
  c.tsc = cov == 3 ? (double)((long double)phi / sqrtl(ln2))
                   : (double)((long double)phi / ln2);
  if (cov == 2) {
    const int nl = (int)(nu + 0.5);
    const double mu = nu - nl;
    c.nl = nl;
    c.mu = mu;
    const double rp = 1.0 / std::tgamma(1.0 + mu);
    const double rm = 1.0 / std::tgamma(1.0 - mu);
    c.rgp = rp;
    c.rgm = rm;
    c.gp = std::tgamma(1.0 + mu);
    c.gm = std::tgamma(1.0 - mu);
    const double pm = 3.14159265358979323846 * mu;
    c.fa = std::fabs(pm) < 1e-300 ? 1.0 : pm / std::sin(pm);
    c.g2 = 0.5 * (rm + rp);
    if (std::fabs(mu) < 0.1) {
      const double u = mu * mu;
      double g = kRg[6];
      for (int k = 5; k >= 0; k--) g = g * u + kRg[k];
      c.g1 = -g;
    } else {
      c.g1 = (rm - rp) / (2.0 * mu);
    }
    c.mnorm = s2 / (std::pow(2.0, nu - 1.0) * std::tgamma(nu));
  }
  return c;
}

// ---------------------------------------------------------------------------
// argument checks
// ---------------------------------------------------------------------------
long long n_indx(int n, int m) {
  // src/rNNGP.cpp:218
  return (long long)((double)(1 + m) / 2 * m + (double)(n - m - 1) * m);
}

int check_neighbors(int n, int m, const int *nnIndx, const int *lu) {
  if (!nnIndx || !lu) return set_err(NNGP_ERR_ARG, "nnIndx/nnIndxLU is NULL");
  const long long ni = n_indx(n, m);
  for (int i = 0; i < n; i++) {
    const long long off = lu[i], cnt = lu[(size_t)n + i];
    if (cnt < 0 || cnt > m || off < 0 || off + cnt > ni)
      return set_err(NNGP_ERR_ARG,
                     "malformed nnIndxLU at site %d (offset %lld, count %lld)",
                     i, off, cnt);
    for (long long j = 0; j < cnt; j++) {
      const int id = nnIndx[off + j];
      if (id < 0 || id >= n)
        return set_err(NNGP_ERR_ARG,
                       "nnIndx entry %lld of site %d is %d, outside [0, n)",
                       j, i, id);
    }
  }
  return NNGP_OK;
}


// ---------------------------------------------------------------------------
// built-in generator
// ---------------------------------------------------------------------------
static uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}
BuiltinRng::BuiltinRng(uint64_t seed) {
  uint64_t x = seed;
  for (int k = 0; k < 4; k++) s[k] = splitmix64(x);
}
uint64_t BuiltinRng::next() {
  const uint64_t r = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return r;
}
double BuiltinRng::unif() {
  return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
double BuiltinRng::norm() {
  if (have) {
    have = 0;
    return cached;
  }
  double u, v, r;
  do {
    u = 2.0 * unif() - 1.0;
    v = 2.0 * unif() - 1.0;
    r = u * u + v * v;
  } while (r >= 1.0 || r == 0.0);
  const double f = std::sqrt(-2.0 * std::log(r) / r);
  cached = v * f;
  have = 1;
  return u * f;
}

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------
__global__ void k_fill_int(int *p, int v) { *p = v; }

int fill_int_launch(int *p, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, p, v);
  return (int)hipGetLastError();
}

int Model::create(int n_, int m_, int p_, const double *coords,
                  const int *nnI, const int *LU, const double *X,
                  const double *y, int cov_, const nngp_opts *opts) {
  n = n_;
  m = m_;
  p = p_;
  cov = cov_;
  precision = opts ? opts->precision : 0;
  nz = (y && (X || p == 0)) ? p + 1 : 0;  // p == 0: y is the one column
  zs = (nz + 1) & ~1;
  nIndx = n_indx(n, m);
  NNGP_TRY(st.create());
  {
    DevBuf tmp;
    NNGP_TRY(tmp.alloc(sizeof(double) * 2 * (size_t)n));
    NNGP_HIP(hipMemcpyAsync(tmp.p, coords, sizeof(double) * 2 * (size_t)n,
                            hipMemcpyHostToDevice, st.s));
    NNGP_TRY(xy.alloc(sizeof(double2) * (size_t)n));
    if (interleave_launch(tmp.as<double>(), n, xy.as<double2>(), st.s))
      return set_err(NNGP_ERR_HIP, "coordinate upload kernel failed");
    NNGP_HIP(hipStreamSynchronize(st.s));
  }
  // the index array carries 64 zero entries of padding: the per-lane kernel
  // reads a full block of M <= 16 indices at every offset
  NNGP_TRY(nnIndx.alloc(sizeof(int) * ((size_t)nIndx + 64)));
  NNGP_HIP(hipMemsetAsync(nnIndx.p, 0, nnIndx.bytes, st.s));
  NNGP_HIP(hipMemcpyAsync(nnIndx.p, nnI, sizeof(int) * (size_t)nIndx,
                          hipMemcpyHostToDevice, st.s));
  lu_closed = true;
  for (int i = 0; i < n && lu_closed; i++) {
    int off, cnt;
    nn_closed(i, m, off, cnt);
    lu_closed = LU[i] == off && LU[(size_t)n + i] == cnt;
  }
  NNGP_TRY(lu.alloc(sizeof(int) * 2 * (size_t)n));
  NNGP_HIP(hipMemcpyAsync(lu.p, LU, sizeof(int) * 2 * (size_t)n,
                          hipMemcpyHostToDevice, st.s));
  if (nz) {
    std::vector<double> hz((size_t)n * zs, 0.0);
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < p; j++) hz[(size_t)i * zs + j] = X[(size_t)j * n + i];
      hz[(size_t)i * zs + p] = y[i];
    }
    NNGP_TRY(Z.alloc(sizeof(double) * hz.size()));
    NNGP_HIP(hipMemcpyAsync(Z.p, hz.data(), sizeof(double) * hz.size(),
                            hipMemcpyHostToDevice, st.s));
    NNGP_HIP(hipStreamSynchronize(st.s));
  }
  for (int s = 0; s < 2; s++) NNGP_TRY(F[s].alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(rows.alloc(sizeof(double) * (size_t)rows_capacity(n) * kStatsMax));
  NNGP_TRY(sums.alloc(sizeof(double) * kSumsLen));
  NNGP_TRY(err.alloc(sizeof(int)));
  if (fill_int_launch(err.as<int>(), INT_MAX, st.s))
    return set_err(NNGP_ERR_HIP, "flag reset kernel failed");
  NNGP_TRY(hsums.alloc(sizeof(double) * (kSumsLen + 1)));
  {
    void *dp = nullptr;
    if (hipHostGetDevicePointer(&dp, hsums.p, 0) == hipSuccess && dp)
      sums_mapped = static_cast<double *>(dp);
    else
      (void)hipGetLastError();  // fall back to a copy per pass
  }
  if (sums_mapped) {
    NNGP_TRY(sig_done.alloc(sizeof(int)));
    if (fill_int_launch(sig_done.as<int>(), 0, st.s))
      return set_err(NNGP_ERR_HIP, "counter reset kernel failed");
    *reinterpret_cast<volatile unsigned *>(hsums.as<double>() + kSumsLen) = 0u;
  }
  NNGP_HIP(hipStreamSynchronize(st.s));
  if (opts && opts->shape[1] > 0)
    NNGP_TRY(pin_shape(opts->shape[0], opts->shape[1], opts->shape[2]));
  return NNGP_OK;
}

int Model::bf(const CovArgs &c, int slot, bool storeB) {
  if (storeB && !B[slot].p)
    // 64 entries of padding: the latent sweep reads chunks of 16
    NNGP_TRY(B[slot].alloc(sizeof(double) * ((size_t)std::max(nIndx, 1LL) + 64)));
  BfArgs a;
  a.xy = xy.as<double2>();
  a.nnIndx = nnIndx.as<int>();
  a.lu = lu.as<int>();
  a.n = n;
  a.m = m;
  a.B = storeB ? B[slot].as<double>() : nullptr;
  a.F = F[slot].as<double>();
  a.Z = nz ? Z.as<double>() : nullptr;
  a.nz = nz;
  a.zs = zs;
  a.err = err.as<int>();  // INT_MAX here: every sums pass resets it
  a.rows = rows.as<double>();
  a.rows_cap = rows_capacity(n);
  a.c = c;
  a.variant = cov_variant(c.cov, c.nu);
  a.mixed = precision == NNGP_MIXED ? 1 : 0;
  a.lu_closed = lu_closed ? 1 : 0;
  BfConfig cfg = bf_default(m, a.variant, !a.mixed);
  if (a.mixed) {  // untuned mixed passes keep the covariance copy in
                  // registers (the fastest per-lane mixed shape at m = 15)
    const BfConfig copy{0, 64, 3};
    if (bf_config_ok(copy, m, a.variant)) cfg = copy;
  }
  if (have_tuned && bf_config_ok(tuned_cfg, m, a.variant)) cfg = tuned_cfg;
  last_bf = a;
  last_cfg = cfg;
  prof_sample = prof && prof_seq++ % prof_every == prof_every - 1;
  const int ev = prof_sample ? ev_begin(0) : -1;
  const int rc = bf_launch(a, cfg, 1, st.s);
  if (ev >= 0) ev_end(ev);
  if (rc == -1)
    return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  if (rc) return set_err(NNGP_ERR_HIP, "updateBF launch failed: %s",
                         hipGetErrorString((hipError_t)rc));
  return NNGP_OK;
}

int Model::pin_shape(int family, int threads, int mode) {
  const BfConfig cfg{family, threads, mode};
  // a matern model takes the per-lane shapes at the closed-form nu (0.5,
  // 1.5, 2.5); passes at any other nu use the wavefront kernel's default
  const int variant = cov == 2 ? kCovMatern15 : cov;
  const bool mixed = precision == NNGP_MIXED;
  if (!bf_config_ok(cfg, m, variant) ||
      (family == 0 && (mixed ? mode == 2 : (mode == 1 || mode == 3))) ||
      (family == 2 && mixed))  // the row kernel is FP64 only
    return set_err(NNGP_ERR_ARG,
                   "launch shape (%d, %d, %d) does not apply to this model",
                   family, threads, mode);
  tuned_cfg = cfg;
  have_tuned = true;
  return NNGP_OK;
}

int Model::autotune(const CovArgs &c, BfTuneReport *rep) {
  BfArgs a;
  a.xy = xy.as<double2>();
  a.nnIndx = nnIndx.as<int>();
  a.lu = lu.as<int>();
  a.n = n;
  a.m = m;
  a.B = nullptr;
  a.F = F[0].as<double>();
  a.Z = nz ? Z.as<double>() : nullptr;
  a.nz = nz;
  a.zs = zs;
  a.err = err.as<int>();
  a.rows = rows.as<double>();
  a.rows_cap = rows_capacity(n);
  a.c = c;
  a.variant = cov_variant(c.cov, c.nu);
  a.mixed = precision == NNGP_MIXED ? 1 : 0;
  a.lu_closed = lu_closed ? 1 : 0;
  BfConfig best;
  const int rc = bf_autotune(a, 1, st.s, &best, rep);
  if (rc) return set_err(NNGP_ERR_HIP, "autotune launch failed (%d)", rc);
  // the timing passes may have flagged a failure: start clean
  if (fill_int_launch(err.as<int>(), INT_MAX, st.s))
    return set_err(NNGP_ERR_HIP, "flag reset kernel failed");
  NNGP_HIP(hipStreamSynchronize(st.s));
  tuned_cfg = best;
  have_tuned = true;
  return NNGP_OK;
}

int Model::sums_launch(int slot, int off) {
  double *base = sums_mapped ? sums_mapped : sums.as<double>();
  double *out = base + off;
  double *err_out = base + 2 * kStatsMax + (off ? 1 : 0);
  const int ev = prof_sample ? ev_begin(1) : -1;
  (void)slot;
  SumsSignal sig;
  if (sums_mapped) {
    sig.done = sig_done.as<int>();
    sig.flag = reinterpret_cast<unsigned *>(sums_mapped + kSumsLen);
    sig.seq = ++sig_seq;
  }
  const int rc = bf_sums_launch(last_bf, last_cfg, 1, out, kStatsMax, err_out,
                                st.s, sums_mapped ? &sig : nullptr);
  if (ev >= 0) ev_end(ev);
  if (rc) return set_err(NNGP_ERR_HIP, "sum kernel launch failed");
  return NNGP_OK;
}

double gram_at(const double *raw, int a, int b) {
  if (a < b) std::swap(a, b);
  return raw[2 + a * (a + 1) / 2 + b];
}

double quad_resid(const PassOut &po, const double *beta, int p) {
  // sum_i (beta . e_X(i) - e_y(i))^2 / F_i from the cross-products
  double q = gram_at(po.gram, p, p);
  for (int a = 0; a < p; a++) {
    double ga = -2.0 * gram_at(po.gram, p, a);
    for (int b = 0; b < p; b++) ga += beta[b] * gram_at(po.gram, a, b);
    q += beta[a] * ga;
  }
  return q;
}

int Model::fetch_sums() {
  if (!sums_mapped) {
    NNGP_HIP(hipMemcpyAsync(hsums.p, sums.p, sizeof(double) * kSumsLen,
                            hipMemcpyDeviceToHost, st.s));
    NNGP_HIP(hipStreamSynchronize(st.s));
  } else {
    // Poll the last sum launch's completion word: about 5 us sooner than
    // the stream's own completion signal.  The spin is bounded (0.5 s);
    // past that, or when the sums never come (a failed launch), the stream
    // synchronization below waits and reports as before.
    const volatile unsigned *flag =
        reinterpret_cast<const volatile unsigned *>(hsums.as<double>() +
                                                    kSumsLen);
    const auto t0 = std::chrono::steady_clock::now();
    bool seen = false;
    for (long spin = 0;; spin++) {
      if (*flag == sig_seq) {
        seen = true;
        break;
      }
      if ((spin & 1023) == 1023 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(500))
        break;
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    if (!seen) NNGP_HIP(hipStreamSynchronize(st.s));
  }
  // kernel timing: collected every 64 launches (and by nngp_model_timing),
  // so that waiting for the events does not undo the polling above
  if (prof && (!sums_mapped || nev >= 64)) NNGP_TRY(collect());
  return NNGP_OK;
}

int Model::ev_begin(int kind) {
  if (nev == (int)evs.size()) {
    EvPair e;
    if (hipEventCreate(&e.a) != hipSuccess) return -1;
    if (hipEventCreate(&e.b) != hipSuccess) {
      (void)hipEventDestroy(e.a);
      return -1;
    }
    evs.push_back(e);
  }
  EvPair &e = evs[nev];
  e.kind = kind;
  if (hipEventRecord(e.a, st.s) != hipSuccess) return -1;
  return nev++;
}

void Model::ev_end(int idx) { (void)hipEventRecord(evs[idx].b, st.s); }

int Model::collect() {
  for (int k = 0; k < nev; k++) {
    float ms = 0.0f;
    NNGP_HIP(hipEventSynchronize(evs[k].b));
    NNGP_HIP(hipEventElapsedTime(&ms, evs[k].a, evs[k].b));
    tms[evs[k].kind] += ms;
    tcount[evs[k].kind] += 1;
  }
  nev = 0;
  return NNGP_OK;
}

Model::~Model() {
  for (auto &e : evs) {
    if (e.a) (void)hipEventDestroy(e.a);
    if (e.b) (void)hipEventDestroy(e.b);
  }
}

int Model::check_err(const char *what, int which) {
  const double bad = hsums.as<double>()[2 * kStatsMax + which];
  if (bad != (double)INT_MAX)
    return set_err(NNGP_ERR_NOT_PD,
                   "c++ Rf_error: dpotrf failed\n(%s: the neighbor covariance"
                   " of site %.0f is not positive definite)",
                   what, bad);
  return NNGP_OK;
}

void Model::unpack(const double *raw, PassOut *out) const {
  out->logDet = raw[0];
  out->q = raw[1];
  for (int k = 0; k < kStatsMax; k++) out->gram[k] = raw[k];
}

int Model::pass(const CovArgs &c, int slot, bool storeB, PassOut *out) {
  NNGP_TRY(bf(c, slot, storeB));
  NNGP_TRY(sums_launch(slot, 0));
  NNGP_TRY(fetch_sums());
  NNGP_TRY(check_err("updateBF", 0));
  unpack(hsums.as<double>(), &slot_out[slot]);
  if (out) *out = slot_out[slot];
  return NNGP_OK;
}

}  // namespace nngp
