// nngp_api.hip — the extern "C" entry points of libnngp_hip.so (declared in
// include/nngp.h).  The MCMC, grid and kriging drivers keep the reference's
// host-side control flow (priors, proposals, accept/reject, RNG order) and
// hand every O(n) loop to the device kernels.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <mutex>
#include <numeric>

#include "nngp_host.hpp"
#include "nngp_strips.hpp"

namespace nngp {
const char *last_error();

static int check_cov(int covModel) {
  if (covModel < 0 || covModel > 3)
    return set_err(NNGP_ERR_COV_MODEL,
                   "c++ Rf_error: cov.model is not correctly specified");
  return NNGP_OK;
}

static int check_nm(int n, int m) {
  if (n < 2) return set_err(NNGP_ERR_ARG, "n must be at least 2 (got %d)", n);
  if (m < 1 || m >= n)
    return set_err(NNGP_ERR_ARG, "n.neighbors must be in [1, n-1] (got %d)",
                   m);
  if (m > 64)
    return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  return NNGP_OK;
}

// src/util.cpp:43-49
static double logit_ab(double t, double a, double b) {
  return std::log((t - a) / (b - t));
}
static double logit_inv(double z, double a, double b) {
  return b - (b - a) / (1 + std::exp(z));
}

// ---------------------------------------------------------------------------
// neighbor search (strip index: nngp_strips.hpp)
// ---------------------------------------------------------------------------
// mode 0: the code-book search (src/nn.cpp:214-257), mode 1: the scan
// (src/nn.cpp:63-111); tied distances are resolved in each one's own
// candidate order (knn_build).
static int knn_pred(int n, int m, const double *coords, int *nnIndx,
                    double *nnDist, int *nnIndxLU, int mode) {
  NNGP_TRY(check_nm(n, m));
  if (!coords || !nnIndx || !nnDist || !nnIndxLU)
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  NNGP_TRY(use_device(nullptr));
  Stream st;
  NNGP_TRY(st.create());
  HostStrips h;
  build_strips(n, coords, coords + n, m, h);
  DevStrips ds;
  NNGP_TRY(ds.upload(h, st.s));
  DevBuf xy, dI, dD, dL;
  NNGP_TRY(upload_xy(coords, n, xy, st.s));
  const long long ni = n_indx(n, m);
  NNGP_TRY(dI.alloc(sizeof(int) * (size_t)ni));
  NNGP_TRY(dD.alloc(sizeof(double) * (size_t)ni));
  NNGP_TRY(dL.alloc(sizeof(int) * 2 * (size_t)n));
  NNGP_TRY(knn_build(coords, n, m, mode, h, ds, xy.as<double2>(), dI.as<int>(),
                     dD.as<double>(), dL.as<int>(), st.s, nullptr));
  NNGP_HIP(hipMemcpyAsync(nnIndx, dI.p, dI.bytes, hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(nnDist, dD.p, dD.bytes, hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(nnIndxLU, dL.p, dL.bytes, hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipStreamSynchronize(st.s));
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// response-model Metropolis-within-Gibbs on a device-resident model
// (src/rNNGP.cpp:192-388)
// ---------------------------------------------------------------------------
struct McmcSpec {
  const double *sigmaSqIG, *tauSqIG, *phiUnif, *nuUnif, *betaStarting;
  double start[4];   // sigma.sq, tau.sq, phi, nu
  double tuning[4];
  int nSamples, verbose, nReport;
  int nRep;
  const int *repIndx;
  // host copies for fit.rep only (src/rNNGP.cpp:299-313)
  const double *X;
  const int *nnIndx, *nnIndxLU;
  double *pBetaSamples, *pThetaSamples, *yRepSamples;
  int *nAccept;
};

static int mcmc_loop(Model &md, const McmcSpec &S, Rng &R,
                     const nngp_opts *opts) {
  const int n = md.n, p = md.p, covModel = md.cov;
  const bool matern = covModel == 2;
  const int nTheta = matern ? 4 : 3;
  const int iS = 0, iT = 1, iP = 2, iN = 3;
  double theta[4] = {S.start[0], S.start[1], S.start[2],
                     matern ? S.start[3] : 0.0};
  const double tuning[4] = {S.tuning[0], S.tuning[1], S.tuning[2],
                            matern ? S.tuning[3] : 0.0};
  double thetaCand[4] = {0, 0, 0, 0};
  const double sA = S.sigmaSqIG[0], sB = S.sigmaSqIG[1];
  const double tA = S.tauSqIG[0], tB = S.tauSqIG[1];
  const double pA = S.phiUnif[0], pB = S.phiUnif[1];
  const double nA = matern ? S.nuUnif[0] : 0.0, nB = matern ? S.nuUnif[1] : 0.0;
  std::vector<double> beta(S.betaStarting, S.betaStarting + p);

  auto cov_of = [&](const double *th) {
    return make_cov(covModel, th[iS], th[iT], th[iP], matern ? th[iN] : 0.0);
  };
  auto log_post = [&](const double *th, double logDet, double Q) {
    // src/rNNGP.cpp:317-324
    double lp = -0.5 * logDet - 0.5 * Q;
    lp += std::log(th[iP] - pA) + std::log(pB - th[iP]);
    lp += -1.0 * (1.0 + sA) * std::log(th[iS]) - sB / th[iS] + std::log(th[iS]);
    lp += -1.0 * (1.0 + tA) * std::log(th[iT]) - tB / th[iT] + std::log(th[iT]);
    if (matern) lp += std::log(th[iN] - nA) + std::log(nB - th[iN]);
    return lp;
  };

  int cur = 0;
  PassOut pc;  // current slot's logDet and cross-products
  NNGP_TRY(md.pass(cov_of(theta), cur, false, &pc));
  std::vector<double> tmp_pp((size_t)p * p, 0.0), tmp_p(p), tmp_p2(p);
  RepDrawer rep;  // fit.rep: the replicate draw on the device
  if (S.nRep > 0) NNGP_TRY(rep.init(n, S.nnIndx, S.nnIndxLU));
  bool thetaUpdate = true;
  int accept = 0, batchAccept = 0, status = 0, repCnt = 0;
  std::chrono::steady_clock::time_point loop_t0;
  if (md.prof) {
    NNGP_HIP(hipStreamSynchronize(md.st.s));
    loop_t0 = std::chrono::steady_clock::now();
  }

  for (int s = 0; s < S.nSamples; s++) {
    if (thetaUpdate) {
      thetaUpdate = false;
      // X^T V^-1 y and X^T V^-1 X of the current B, F (src/rNNGP.cpp:281-291)
      for (int i = 0; i < p; i++) {
        tmp_p[i] = gram_at(pc.gram, p, i);
        for (int j = 0; j <= i; j++) tmp_pp[j * p + i] = gram_at(pc.gram, i, j);
      }
      if (h_potrf_L(p, tmp_pp.data()) || h_potri_L(p, tmp_pp.data()))
        return set_err(NNGP_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
      h_symv_L(p, tmp_pp.data(), tmp_p.data(), tmp_p2.data());
      if (h_potrf_L(p, tmp_pp.data()))
        return set_err(NNGP_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
    }
    // beta ~ N(tmp_p2, tmp_pp tmp_pp^T) (src/util.cpp:28-41)
    for (int i = 0; i < p; i++) beta[i] = R.rnorm(0, 1);
    h_trmv_L(p, tmp_pp.data(), beta.data());
    for (int i = 0; i < p; i++) beta[i] += tmp_p2[i];

    if (S.nRep && S.repIndx[s]) {
      // fitted-value replicate: X beta + an NNGP draw with the current B, F
      // (src/rNNGP.cpp:299-313, updateRep at :24-42)
      double *yr = S.yRepSamples + (size_t)repCnt * n;
      for (int i = 0; i < n; i++) yr[i] = 0.0;
      for (int j = 0; j < p; j++)
        for (int i = 0; i < n; i++) yr[i] += beta[j] * S.X[(size_t)j * n + i];
      // B of the current theta (the slot's passes do not store it), the n
      // normals in site order from the chain's stream, then the draw on the
      // device; only the n-vector comes back
      NNGP_TRY(md.pass(cov_of(theta), cur, true, nullptr));
      double *z = rep.z();
      for (int i = 0; i < n; i++) z[i] = R.rnorm(0.0, 1.0);
      NNGP_TRY(rep.draw(md.B[cur].as<double>(), md.F[cur].as<double>(),
                        md.nnIndx.as<int>(), md.lu.as<int>(), md.st.s));
      const double *w = rep.v();
      for (int i = 0; i < n; i++) yr[i] += w[i];
      repCnt++;
    }

    // proposals (src/rNNGP.cpp:327-333), drawn before the device work so
    // that both passes go out back to back; the RNG order is unchanged
    thetaCand[iP] =
        logit_inv(R.rnorm(logit_ab(theta[iP], pA, pB), tuning[iP]), pA, pB);
    thetaCand[iS] = std::exp(R.rnorm(std::log(theta[iS]), tuning[iS]));
    thetaCand[iT] = std::exp(R.rnorm(std::log(theta[iT]), tuning[iT]));
    if (matern)
      thetaCand[iN] =
          logit_inv(R.rnorm(logit_ab(theta[iN], nA, nB), tuning[iN]), nA, nB);

    // the candidate's B, F, log-det and cross-products in one pass; both
    // quadratic forms Q(r, r), r = X beta - y, follow from cross-products
    // (src/rNNGP.cpp:299-315, 336-338)
    PassOut pn;
    NNGP_TRY(md.pass(cov_of(thetaCand), 1 - cur, false, &pn));
    const double QCurrent = quad_resid(pc, beta.data(), p);
    const double lpCur = log_post(theta, pc.logDet, QCurrent);
    const double QCand = quad_resid(pn, beta.data(), p);
    const double lpCand = log_post(thetaCand, pn.logDet, QCand);

    if (R.runif(0.0, 1.0) <= std::exp(lpCand - lpCur)) {
      thetaUpdate = true;
      for (int k = 0; k < nTheta; k++) theta[k] = thetaCand[k];
      cur = 1 - cur;
      pc = pn;
      accept++;
      batchAccept++;
    }
    for (int k = 0; k < p; k++) S.pBetaSamples[(size_t)s * p + k] = beta[k];
    for (int k = 0; k < nTheta; k++)
      S.pThetaSamples[(size_t)s * nTheta + k] = theta[k];

    if (status == S.nReport) {
      if (S.verbose) {
        printf("Sampled: %i of %i, %3.2f%%\n", s, S.nSamples,
               100.0 * s / S.nSamples);
        printf("Report interval Metrop. Acceptance rate: %3.2f%%\n",
               100.0 * batchAccept / S.nReport);
        printf("Overall Metrop. Acceptance rate: %3.2f%%\n",
               100.0 * accept / s);
        printf("-------------------------------------------------\n");
        fflush(stdout);
      }
      batchAccept = 0;
      status = 0;
    }
    status++;
    if (S.nAccept) *S.nAccept = accept;
    if (opts && opts->interrupt && opts->interrupt(opts->interrupt_arg))
      return set_err(NNGP_ERR_INTERRUPT, "interrupted after %d samples", s + 1);
  }
  if (md.prof) {
    NNGP_HIP(hipStreamSynchronize(md.st.s));
    md.loop_ms += std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - loop_t0)
                      .count();
    md.loop_iters += S.nSamples;
  }
  if (S.verbose) {
    printf("Sampled: %i of %i, %3.2f%%\n", S.nSamples, S.nSamples, 100.0);
    printf("Overall Metrop. Acceptance rate: %3.2f%%\n",
           S.nSamples ? 100.0 * accept / S.nSamples : 0.0);
    fflush(stdout);
  }
  if (S.nAccept) *S.nAccept = accept;
  return NNGP_OK;
}

static void print_model(int n, int p, int m, int covModel, int nSamples,
                        const double *sIG, const double *tIG,
                        const double *phiU, const double *nuU, int device) {
  static const char *names[] = {"exponential", "spherical", "matern",
                                "gaussian"};
  printf("----------------------------------------\n");
  printf("\tModel description\n");
  printf("----------------------------------------\n");
  printf("NNGP Response model fit with %i observations.\n\n", n);
  printf("Number of covariates %i (including intercept if specified).\n\n", p);
  printf("Using the %s spatial correlation model.\n\n", names[covModel]);
  printf("Using %i nearest neighbors.\n\n", m);
  printf("Number of MCMC samples %i.\n\n", nSamples);
  printf("Priors and hyperpriors:\n\tbeta flat.\n");
  printf("\tsigma.sq IG hyperpriors shape=%.5f and scale=%.5f\n", sIG[0], sIG[1]);
  printf("\ttau.sq IG hyperpriors shape=%.5f and scale=%.5f\n", tIG[0], tIG[1]);
  printf("\tphi Unif hyperpriors a=%.5f and b=%.5f\n", phiU[0], phiU[1]);
  if (covModel == 2)
    printf("\tnu Unif hyperpriors a=%.5f and b=%.5f\n", nuU[0], nuU[1]);
  printf("\nComputed on HIP device %d.\n", device);
  fflush(stdout);
}

// ---------------------------------------------------------------------------
// Q on device buffers without allocating
// ---------------------------------------------------------------------------
// k_quadform's per-block partial sums go to this file-scope array instead of
// a buffer allocated per call.  Every device holds its own copy, at its own
// address: that address is looked up once per device (hipGetSymbolAddress,
// with the device current) and only it is handed to kernels.  The array is
// shared by every call on a device and keeps the last call's partial sums;
// each launch writes every partial it then reads, so nothing needs clearing.
static __device__ __align__(256) double scratch[kQPartialRows];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices] = {};

static int scratch_on_current_device(double **out) {
  int dev = 0;
  NNGP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_err(NNGP_ERR_UNSUPPORTED, "device ordinal %d above %d", dev,
                   kMaxDevices - 1);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    NNGP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return NNGP_OK;
}

}  // namespace nngp

using namespace nngp;

extern "C" {

const char *nngp_last_error(void) { return last_error(); }

int nngp_version(void) { return 10000; }

int nngp_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

// ---------------------------------------------------------------------------
// neighbor search
// ---------------------------------------------------------------------------
int nngp_mkNNIndxCB(int n, int m, const double *coords, int *nnIndx,
                    double *nnDist, int *nnIndxLU, int nThreads) {
  (void)nThreads;
  clear_err();
  return knn_pred(n, m, coords, nnIndx, nnDist, nnIndxLU, 0);
}

int nngp_mkNNIndx(int n, int m, const double *coords, int *nnIndx,
                  double *nnDist, int *nnIndxLU, int nThreads) {
  (void)nThreads;
  clear_err();
  return knn_pred(n, m, coords, nnIndx, nnDist, nnIndxLU, 1);
}

int nngp_nn2(int n, const double *coords, int q, const double *coords0,
             int k, int *nnIndx0, double *nnDist0) {
  clear_err();
  if (n < 1 || q < 0 || k < 1 || k > n)
    return set_err(NNGP_ERR_ARG, "nn2 needs 1 <= k <= n (n=%d, k=%d)", n, k);
  if (k > 64) return set_err(NNGP_ERR_UNSUPPORTED, "k = %d exceeds 64", k);
  if (!coords || (q > 0 && (!coords0 || !nnIndx0)))
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  if (q == 0) return NNGP_OK;
  NNGP_TRY(use_device(nullptr));
  Stream st;
  NNGP_TRY(st.create());
  HostStrips h;
  build_strips(n, coords, coords + n, k, h);
  DevStrips ds;
  NNGP_TRY(ds.upload(h, st.s));
  DevBuf q0, dI, dD;
  NNGP_TRY(upload_xy(coords0, q, q0, st.s));
  NNGP_TRY(dI.alloc(sizeof(int) * (size_t)q * k));
  if (nnDist0) NNGP_TRY(dD.alloc(sizeof(double) * (size_t)q * k));
  Nn2Args a;
  a.s = ds.idx;
  a.q = q0.as<double2>();
  a.nq = q;
  a.k = k;
  a.idx = dI.as<int>();
  a.dist = nnDist0 ? dD.as<double>() : nullptr;
  if (nn2_launch(a, st.s)) return set_err(NNGP_ERR_HIP, "nn2 launch failed");
  NNGP_HIP(hipMemcpyAsync(nnIndx0, dI.p, dI.bytes, hipMemcpyDeviceToHost, st.s));
  if (nnDist0)
    NNGP_HIP(hipMemcpyAsync(nnDist0, dD.p, dD.bytes, hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipStreamSynchronize(st.s));
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// kernels with host buffers
// ---------------------------------------------------------------------------
static int bf_host(double *B, double *F, const double *coords,
                   const int *nnIndx, const int *nnIndxLU, int n, int m,
                   const CovArgs &c, double *logDet, const nngp_opts *opts) {
  NNGP_TRY(check_nm(n, m));
  if (!B || !F || !coords) return set_err(NNGP_ERR_ARG, "NULL buffer");
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  NNGP_TRY(use_device(opts));
  Model md;
  NNGP_TRY(md.create(n, m, 0, coords, nnIndx, nnIndxLU, nullptr, nullptr,
                     c.cov, opts));
  PassOut out;
  NNGP_TRY(md.pass(c, 0, true, &out));
  // B for site 0 is never written (it has no neighbors)
  NNGP_HIP(hipMemcpy(B, md.B[0].p, sizeof(double) * (size_t)md.nIndx,
                     hipMemcpyDeviceToHost));
  NNGP_HIP(hipMemcpy(F, md.F[0].p, sizeof(double) * (size_t)n,
                     hipMemcpyDeviceToHost));
  if (logDet) *logDet = out.logDet;
  return NNGP_OK;
}

int nngp_updateBF(double *B, double *F, const double *coords,
                  const int *nnIndx, const int *nnIndxLU, int n, int m,
                  const double *theta, int covModel, double *logDet,
                  const nngp_opts *opts) {
  clear_err();
  NNGP_TRY(check_cov(covModel));
  if (!theta) return set_err(NNGP_ERR_ARG, "theta is NULL");
  const CovArgs c = make_cov(covModel, theta[0], theta[1], theta[2],
                             covModel == 2 ? theta[3] : 0.0);
  return bf_host(B, F, coords, nnIndx, nnIndxLU, n, m, c, logDet, opts);
}

int nngp_updateConjBF(double *B, double *F, const double *coords,
                      const int *nnIndx, const int *nnIndxLU, int n, int m,
                      double phi, double alpha, double nu, int covModel,
                      const nngp_opts *opts) {
  clear_err();
  NNGP_TRY(check_cov(covModel));
  const CovArgs c = make_cov(covModel, 1.0, alpha, phi, nu);
  return bf_host(B, F, coords, nnIndx, nnIndxLU, n, m, c, nullptr, opts);
}

int nngp_Q(const double *B, const double *F, const double *u,
           const double *v, int n, int m, const int *nnIndx,
           const int *nnIndxLU, double *q) {
  clear_err();
  NNGP_TRY(check_nm(n, m));
  if (!B || !F || !u || !v || !q) return set_err(NNGP_ERR_ARG, "NULL buffer");
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  NNGP_TRY(use_device(nullptr));
  Stream st;
  NNGP_TRY(st.create());
  const long long ni = n_indx(n, m);
  DevBuf dB, dF, du, dv, dI, dL, part, out;
  NNGP_TRY(dB.alloc(sizeof(double) * (size_t)ni));
  NNGP_TRY(dF.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(du.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(dv.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(dI.alloc(sizeof(int) * (size_t)ni));
  NNGP_TRY(dL.alloc(sizeof(int) * 2 * (size_t)n));
  NNGP_TRY(part.alloc(sizeof(double) * (size_t)rows_capacity(n)));
  NNGP_TRY(out.alloc(sizeof(double)));
  NNGP_HIP(hipMemcpyAsync(dB.p, B, dB.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dF.p, F, dF.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(du.p, u, du.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dv.p, v, dv.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dI.p, nnIndx, dI.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dL.p, nnIndxLU, dL.bytes, hipMemcpyHostToDevice, st.s));
  QArgs a{dB.as<double>(), dF.as<double>(), du.as<double>(), dv.as<double>(),
          dI.as<int>(), dL.as<int>(), n, part.as<double>(), out.as<double>()};
  if (q_launch(a, st.s)) return set_err(NNGP_ERR_HIP, "Q launch failed");
  NNGP_HIP(hipMemcpyAsync(q, out.p, sizeof(double), hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipStreamSynchronize(st.s));
  return NNGP_OK;
}

// Q on device buffers without allocating (scratch: nngp_api.hip's
// file-scope array, see scratch_on_current_device)
int nngp_Q_noalloc(const double *B, const double *F, const double *u,
                   const double *v, int n, int m, const int *nnIndx,
                   const int *nnIndxLU, double *q, void *stream) {
  clear_err();
  NNGP_TRY(check_nm(n, m));
  if (!B || !F || !u || !v || !nnIndx || !nnIndxLU || !q)
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  if (n_indx(n, m) > (long long)INT_MAX)
    return set_err(NNGP_ERR_ARG, "n = %d, m = %d: nIndx exceeds 2^31 - 1", n,
                   m);
  const int rows = q_partial_rows(n);
  if (rows > kQPartialRows)
    return set_err(NNGP_ERR_UNSUPPORTED,
                   "n = %d needs %d partial sums; the scratch holds %d", n,
                   rows, kQPartialRows);
  double *part = nullptr;
  NNGP_TRY(scratch_on_current_device(&part));
  QArgs a{B, F, u, v, nnIndx, nnIndxLU, n, part, q};
  if (q_launch(a, static_cast<hipStream_t>(stream)))
    return set_err(NNGP_ERR_HIP, "Q launch failed");
  return NNGP_OK;
}

// src/rNNGP.cpp:24-42 with the normals supplied
int nngp_prior_draw(const double *B, const double *F, const int *nnIndx,
                    const int *nnIndxLU, int n, const double *z, double *w) {
  clear_err();
  if (n < 1 || !B || !F || !nnIndx || !nnIndxLU || !z || !w)
    return set_err(NNGP_ERR_ARG, "NULL buffer or n < 1");
  for (int i = 0; i < n; i++) {
    double acc = 0.0;
    const int off = nnIndxLU[i], cnt = nnIndxLU[(size_t)n + i];
    for (int j = 0; j < cnt; j++) {
      const int id = nnIndx[off + j];
      if (id < 0 || id >= i)
        return set_err(NNGP_ERR_ARG,
                       "site %d has neighbor %d that is not a predecessor", i,
                       id);
      acc += B[off + j] * w[id];
    }
    w[i] = acc + std::sqrt(F[i]) * z[i];
  }
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// rNNGP (src/rNNGP.cpp:105-420)
// ---------------------------------------------------------------------------
int nngp_rNNGP(const double *y, const double *X, int p, int n, int m,
               const double *coords, int covModel, const int *nnIndx,
               const int *nnIndxLU, const double *sigmaSqIG,
               const double *tauSqIG, const double *phiUnif,
               const double *nuUnif, const double *betaStarting,
               double sigmaSqStarting, double tauSqStarting,
               double phiStarting, double nuStarting, double sigmaSqTuning,
               double tauSqTuning, double phiTuning, double nuTuning,
               int nSamples, int nThreads, int verbose, int nReport,
               int nRep, const int *repIndx, const nngp_rng *rng,
               const nngp_opts *opts, double *pBetaSamples,
               double *pThetaSamples, double *yRepSamples, int *nAccept) {
  (void)nThreads;
  clear_err();
  NNGP_TRY(check_cov(covModel));
  NNGP_TRY(check_nm(n, m));
  if (p < 1 || p > kMaxP)
    return set_err(NNGP_ERR_UNSUPPORTED, "p = %d covariates (1..%d supported)",
                   p, kMaxP);
  if (!y || !X || !coords || !sigmaSqIG || !tauSqIG || !phiUnif ||
      !betaStarting || !pBetaSamples || !pThetaSamples || nSamples < 0)
    return set_err(NNGP_ERR_ARG, "NULL buffer or negative nSamples");
  const bool matern = covModel == 2;
  if (matern && !nuUnif) return set_err(NNGP_ERR_ARG, "nuUnif is NULL");
  if (nRep > 0 && (!repIndx || !yRepSamples))
    return set_err(NNGP_ERR_ARG, "nRep > 0 needs repIndx and yRepSamples");
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  NNGP_TRY(use_device(opts));

  if (verbose)
    print_model(n, p, m, covModel, nSamples, sigmaSqIG, tauSqIG, phiUnif,
                nuUnif, opts ? opts->device : 0);
  Rng R(rng, opts ? opts->seed : 1ull);
  Model md;
  NNGP_TRY(md.create(n, m, p, coords, nnIndx, nnIndxLU, X, y, covModel, opts));
  McmcSpec S;
  S.sigmaSqIG = sigmaSqIG;
  S.tauSqIG = tauSqIG;
  S.phiUnif = phiUnif;
  S.nuUnif = nuUnif;
  S.betaStarting = betaStarting;
  S.start[0] = sigmaSqStarting;
  S.start[1] = tauSqStarting;
  S.start[2] = phiStarting;
  S.start[3] = nuStarting;
  S.tuning[0] = sigmaSqTuning;
  S.tuning[1] = tauSqTuning;
  S.tuning[2] = phiTuning;
  S.tuning[3] = nuTuning;
  S.nSamples = nSamples;
  S.verbose = verbose;
  S.nReport = nReport;
  S.nRep = nRep;
  S.repIndx = repIndx;
  S.X = X;
  S.nnIndx = nnIndx;
  S.nnIndxLU = nnIndxLU;
  S.pBetaSamples = pBetaSamples;
  S.pThetaSamples = pThetaSamples;
  S.yRepSamples = yRepSamples;
  S.nAccept = nAccept;
  return mcmc_loop(md, S, R, opts);
}

// ---------------------------------------------------------------------------
// cNNGP (src/cNNGP.cpp:73-358)
// ---------------------------------------------------------------------------
int nngp_cNNGP(const double *y, const double *X, int p, int n,
               const double *coords, const double *thetaAlpha,
               const int *nnIndx, const int *nnIndxLU, const double *X0,
               const double *coords0, int n0, const int *nnIndx0, int g,
               int m, const double *sigmaSqIG, int covModel, int nThreads,
               int verbose, const nngp_opts *opts, double *betaHat,
               double *ab, double *bBInv, double *bb, double *y0Hat,
               double *y0HatVar) {
  (void)nThreads;
  clear_err();
  NNGP_TRY(check_cov(covModel));
  NNGP_TRY(check_nm(n, m));
  if (p < 1 || p > kMaxP)
    return set_err(NNGP_ERR_UNSUPPORTED, "p = %d covariates (1..%d supported)",
                   p, kMaxP);
  if (!y || !X || !coords || !thetaAlpha || !sigmaSqIG || !betaHat || !ab ||
      !bBInv || !bb || g < 0 || n0 < 0)
    return set_err(NNGP_ERR_ARG, "NULL buffer or negative size");
  if (n0 > 0 && (!X0 || !coords0 || !nnIndx0 || !y0Hat || !y0HatVar))
    return set_err(NNGP_ERR_ARG, "prediction buffers are NULL");
  for (long long k = 0; k < (long long)n0 * m; k++)
    if (nnIndx0[k] < 0 || nnIndx0[k] >= n)
      return set_err(NNGP_ERR_ARG, "nnIndx0 entry %lld outside [0, n)", k);
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  NNGP_TRY(use_device(opts));
  const int nTheta = covModel == 2 ? 3 : 2;
  Model md;
  NNGP_TRY(md.create(n, m, p, coords, nnIndx, nnIndxLU, X, y, covModel, opts));
  DevBuf dX, dy, dxy0, dX0, dnn0, dbeta, dVb, dy0, dy0v, perr;
  NNGP_TRY(perr.alloc(sizeof(int)));
  if (fill_int_launch(perr.as<int>(), INT_MAX, md.st.s))
    return set_err(NNGP_ERR_HIP, "flag reset kernel failed");
  hipStream_t st = md.st.s;
  if (n0 > 0) {
    NNGP_TRY(dX.alloc(sizeof(double) * (size_t)n * p));
    NNGP_TRY(dy.alloc(sizeof(double) * (size_t)n));
    NNGP_TRY(dX0.alloc(sizeof(double) * (size_t)n0 * p));
    NNGP_TRY(dnn0.alloc(sizeof(int) * (size_t)n0 * m));
    NNGP_TRY(dy0.alloc(sizeof(double) * (size_t)n0 * g));
    NNGP_TRY(dy0v.alloc(sizeof(double) * (size_t)n0 * g));
    NNGP_HIP(hipMemcpyAsync(dX.p, X, dX.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(dy.p, y, dy.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(dX0.p, X0, dX0.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(dnn0.p, nnIndx0, dnn0.bytes, hipMemcpyHostToDevice, st));
    NNGP_TRY(upload_xy(coords0, n0, dxy0, st));
  }
  NNGP_TRY(dbeta.alloc(sizeof(double) * kMaxP * (size_t)std::max(g, 1)));
  NNGP_TRY(dVb.alloc(sizeof(double) * kMaxP * kMaxP * (size_t)std::max(g, 1)));
  std::vector<double> tmp_pp((size_t)p * p), tmp_p(p);
  for (int k = 0; k < g; k++) {
    const double phi = thetaAlpha[(size_t)k * nTheta];
    const double alpha = thetaAlpha[(size_t)k * nTheta + 1];
    const double nu = covModel == 2 ? thetaAlpha[(size_t)k * nTheta + 2] : 0.0;
    const CovArgs c = make_cov(covModel, 1.0, alpha, phi, nu);
    PassOut po;
    NNGP_TRY(md.pass(c, 0, false, &po));
    std::fill(tmp_pp.begin(), tmp_pp.end(), 0.0);
    for (int i = 0; i < p; i++) {
      tmp_p[i] = gram_at(po.gram, p, i);
      for (int j = 0; j <= i; j++) tmp_pp[j * p + i] = gram_at(po.gram, i, j);
    }
    if (h_potrf_L(p, tmp_pp.data()) || h_potri_L(p, tmp_pp.data()))
      return set_err(NNGP_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
    double *bk = betaHat + (size_t)k * p;
    h_symv_L(p, tmp_pp.data(), tmp_p.data(), bk);
    const double a = sigmaSqIG[0] + 1.0 * n / 2.0;
    double dot = 0.0;
    for (int i = 0; i < p; i++) dot += bk[i] * tmp_p[i];
    const double b = sigmaSqIG[1] + 0.5 * (gram_at(po.gram, p, p) - dot);
    ab[(size_t)k * 2] = a;
    ab[(size_t)k * 2 + 1] = b;
    for (int i = 0; i < p; i++)
      for (int j = 0; j < p; j++)
        bBInv[(size_t)k * p * p + j * p + i] =
            i >= j ? tmp_pp[j * p + i] : tmp_pp[i * p + j];
    for (int i = 0; i < p; i++) bb[(size_t)k * p + i] = tmp_p[i];
    if (n0 > 0) {
      double *db = dbeta.as<double>() + (size_t)k * kMaxP;
      double *dv = dVb.as<double>() + (size_t)k * kMaxP * kMaxP;
      NNGP_HIP(hipMemcpyAsync(db, bk, sizeof(double) * p, hipMemcpyHostToDevice, st));
      NNGP_HIP(hipMemcpyAsync(dv, tmp_pp.data(), sizeof(double) * p * p,
                              hipMemcpyHostToDevice, st));
      ConjPredArgs pa;
      pa.xy = md.xy.as<double2>();
      pa.X = dX.as<double>();
      pa.y = dy.as<double>();
      pa.n = n;
      pa.p = p;
      pa.m = m;
      pa.xy0 = dxy0.as<double2>();
      pa.X0 = dX0.as<double>();
      pa.n0 = n0;
      pa.nn0 = dnn0.as<int>();
      pa.beta = db;
      pa.Vb = dv;
      pa.a = a;
      pa.b = b;
      pa.alpha = alpha;
      pa.y0 = dy0.as<double>() + (size_t)k * n0;
      pa.y0var = dy0v.as<double>() + (size_t)k * n0;
      pa.err = perr.as<int>();
      pa.c = c;
      if (conj_pred_launch(pa, st))
        return set_err(NNGP_ERR_HIP, "prediction launch failed");
    }
    if (verbose) {
      if (covModel == 2)
        printf("Set phi=%.5f, nu=%.5f, and alpha=%.5f\n", phi, nu, alpha);
      else
        printf("Set phi=%.5f and alpha=%.5f\n", phi, alpha);
      fflush(stdout);
    }
    if (opts && opts->interrupt && opts->interrupt(opts->interrupt_arg))
      return set_err(NNGP_ERR_INTERRUPT, "interrupted after %d of %d sets",
                     k + 1, g);
  }
  if (n0 > 0) {
    int bad = INT_MAX;
    NNGP_HIP(hipMemcpyAsync(y0Hat, dy0.p, dy0.bytes, hipMemcpyDeviceToHost, st));
    NNGP_HIP(hipMemcpyAsync(y0HatVar, dy0v.p, dy0v.bytes, hipMemcpyDeviceToHost, st));
    NNGP_HIP(hipMemcpyAsync(&bad, perr.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NNGP_HIP(hipStreamSynchronize(st));
    if (bad != INT_MAX)
      return set_err(NNGP_ERR_NOT_PD,
                     "c++ Rf_error: dpotrf failed\n(prediction system of "
                     "held-out site %d is not positive definite)", bad);
  }
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// rNNGPPredict (src/rNNGPPredict.cpp:25-222)
// ---------------------------------------------------------------------------
int nngp_rNNGPPredict(const double *X, const double *y, const double *coords,
                      int n, int p, int m, const double *X0,
                      const double *coords0, int q, const int *nnIndx0,
                      const double *betaSamples, const double *thetaSamples,
                      int nSamples, int covModel, int nThreads, int verbose,
                      int nReport, const double *z, const nngp_rng *rng,
                      const nngp_opts *opts, double *pY0) {
  (void)nThreads;
  (void)nReport;
  clear_err();
  NNGP_TRY(check_cov(covModel));
  if (n < 1 || m < 1 || m > n)
    return set_err(NNGP_ERR_ARG, "need 1 <= n.neighbors <= n");
  if (m > 64) return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  if (p < 1 || p > kMaxP)
    return set_err(NNGP_ERR_UNSUPPORTED, "p = %d covariates (1..%d supported)",
                   p, kMaxP);
  if (q < 0 || nSamples < 0) return set_err(NNGP_ERR_ARG, "negative size");
  if (q == 0 || nSamples == 0) return NNGP_OK;
  if (!X || !y || !coords || !X0 || !coords0 || !nnIndx0 || !betaSamples ||
      !thetaSamples || !pY0)
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  for (long long k = 0; k < (long long)q * m; k++)
    if (nnIndx0[k] < 0 || nnIndx0[k] >= n)
      return set_err(NNGP_ERR_ARG, "nnIndx0 entry %lld outside [0, n)", k);
  NNGP_TRY(use_device(opts));
  const int nTheta = covModel == 2 ? 4 : 3;
  if (verbose) {
    printf("----------------------------------------\n");
    printf("\tPrediction description\n");
    printf("----------------------------------------\n");
    printf("NNGP Response model fit with %i observations.\n\n", n);
    printf("Number of covariates %i (including intercept if specified).\n\n", p);
    printf("Using %i nearest neighbors.\n\n", m);
    printf("Number of MCMC samples %i.\n\n", nSamples);
    printf("Predicting at %i locations.\n\n", q);
    fflush(stdout);
  }
  // z: the reference draws q*S normals up front, z[i*S + s] for site i,
  // sample s (single-thread order of src/rNNGPPredict.cpp:131-181)
  std::vector<double> zown;
  if (!z) {
    Rng R(rng, opts ? opts->seed : 1ull);
    zown.resize((size_t)q * nSamples);
    for (auto &v : zown) v = R.rnorm(0.0, 1.0);
    z = zown.data();
  }
  Stream st;
  NNGP_TRY(st.create());
  DevBuf dxy, dX, dy, dxy0, dX0, dnn0, dbeta, dtheta, dz, dy0, dcov, derr;
  NNGP_TRY(upload_xy(coords, n, dxy, st.s));
  NNGP_TRY(upload_xy(coords0, q, dxy0, st.s));
  NNGP_TRY(dX.alloc(sizeof(double) * (size_t)n * p));
  NNGP_TRY(dy.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(dX0.alloc(sizeof(double) * (size_t)q * p));
  NNGP_TRY(dnn0.alloc(sizeof(int) * (size_t)q * m));
  NNGP_TRY(dbeta.alloc(sizeof(double) * (size_t)p * nSamples));
  NNGP_TRY(dtheta.alloc(sizeof(double) * (size_t)nTheta * nSamples));
  NNGP_TRY(dz.alloc(sizeof(double) * (size_t)q * nSamples));
  NNGP_TRY(dy0.alloc(sizeof(double) * (size_t)q * nSamples));
  NNGP_TRY(derr.alloc(sizeof(int)));
  NNGP_HIP(hipMemcpyAsync(dX.p, X, dX.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dy.p, y, dy.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dX0.p, X0, dX0.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dnn0.p, nnIndx0, dnn0.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dbeta.p, betaSamples, dbeta.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dtheta.p, thetaSamples, dtheta.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dz.p, z, dz.bytes, hipMemcpyHostToDevice, st.s));
  std::vector<CovArgs> hc;
  const bool generic = covModel == 2 || m > 16;
  if (generic) {
    hc.resize(nSamples);
    for (int s = 0; s < nSamples; s++) {
      const double *th = thetaSamples + (size_t)s * nTheta;
      hc[s] = make_cov(covModel, th[0], th[1], th[2], covModel == 2 ? th[3] : 0.0);
    }
    NNGP_TRY(dcov.alloc(sizeof(CovArgs) * hc.size()));
    NNGP_HIP(hipMemcpyAsync(dcov.p, hc.data(), dcov.bytes, hipMemcpyHostToDevice, st.s));
  }
  if (fill_int_launch(derr.as<int>(), INT_MAX, st.s))
    return set_err(NNGP_ERR_HIP, "flag reset kernel failed");
  PredArgs a;
  a.xy = dxy.as<double2>();
  a.X = dX.as<double>();
  a.y = dy.as<double>();
  a.n = n;
  a.p = p;
  a.m = m;
  a.xy0 = dxy0.as<double2>();
  a.X0 = dX0.as<double>();
  a.q = q;
  a.nn0 = dnn0.as<int>();
  a.beta = dbeta.as<double>();
  a.theta = dtheta.as<double>();
  a.S = nSamples;
  a.nTheta = nTheta;
  a.cov = covModel;
  a.z = dz.as<double>();
  a.y0 = dy0.as<double>();
  a.err = derr.as<int>();
  a.unit = make_cov(covModel, 1.0, 0.0, 1.0, 0.0);
  a.cov_s = generic ? dcov.as<CovArgs>() : nullptr;
  const int chunk = 65535;
  for (int s0 = 0; s0 < q; s0 += chunk) {
    a.site0 = s0;
    a.nsite = std::min(chunk, q - s0);
    if (pred_launch(a, st.s)) return set_err(NNGP_ERR_HIP, "kriging launch failed");
  }
  int bad = INT_MAX;
  NNGP_HIP(hipMemcpyAsync(pY0, dy0.p, dy0.bytes, hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(&bad, derr.p, sizeof(int), hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipStreamSynchronize(st.s));
  if (bad != INT_MAX)
    return set_err(NNGP_ERR_NOT_PD,
                   "c++ Rf_error: dpotrf failed\n(kriging system of new site %d"
                   " is not positive definite)", bad);
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// sNNGPPredict (src/sNNGPPredict.cpp:25-241), Gaussian family
// ---------------------------------------------------------------------------
int nngp_sNNGPPredict(const double *X, const double *y, const double *coords,
                      int n, int p, int m, const double *X0,
                      const double *coords0, int q, const int *nnIndx0,
                      const double *betaSamples, const double *thetaSamples,
                      const double *wSamples, int nSamples, int family,
                      int covModel, int nThreads, int verbose, int nReport,
                      const double *z, const nngp_rng *rng,
                      const nngp_opts *opts, double *pY0, double *pW0) {
  (void)nThreads;
  (void)nReport;
  (void)y;  // the latent predictor conditions on w, not on y
  clear_err();
  NNGP_TRY(check_cov(covModel));
  if (family != 1)
    return set_err(NNGP_ERR_UNSUPPORTED,
                   "family %d: only the Gaussian family (1) is supported",
                   family);
  if (n < 1 || m < 1 || m > n)
    return set_err(NNGP_ERR_ARG, "need 1 <= n.neighbors <= n");
  if (m > 64) return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  if (p < 1 || p > kMaxP)
    return set_err(NNGP_ERR_UNSUPPORTED, "p = %d covariates (1..%d supported)",
                   p, kMaxP);
  if (q < 0 || nSamples < 0) return set_err(NNGP_ERR_ARG, "negative size");
  if (q == 0 || nSamples == 0) return NNGP_OK;
  if (!X || !coords || !X0 || !coords0 || !nnIndx0 || !betaSamples ||
      !thetaSamples || !wSamples || !pY0 || !pW0)
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  for (long long k = 0; k < (long long)q * m; k++)
    if (nnIndx0[k] < 0 || nnIndx0[k] >= n)
      return set_err(NNGP_ERR_ARG, "nnIndx0 entry %lld outside [0, n)", k);
  NNGP_TRY(use_device(opts));
  const int nTheta = covModel == 2 ? 4 : 3;
  if (verbose) {
    printf("-------------------------------------------------\n");
    printf("\t\tPredicting\n");
    printf("-------------------------------------------------\n");
    fflush(stdout);
  }
  // the reference draws all q*S normals of w0, then all q*S of y0, each
  // site-major (z[i*S + s], the single-thread order of :135-147, 192)
  const size_t qs = (size_t)q * nSamples;
  std::vector<double> zown;
  if (!z) {
    Rng R(rng, opts ? opts->seed : 1ull);
    zown.resize(2 * qs);
    for (auto &v : zown) v = R.rnorm(0.0, 1.0);
    z = zown.data();
  }
  Stream st;
  NNGP_TRY(st.create());
  DevBuf dxy, dX, dxy0, dX0, dnn0, dbeta, dtheta, dw, dz, dy0, dw0, dcov, derr;
  NNGP_TRY(upload_xy(coords, n, dxy, st.s));
  NNGP_TRY(upload_xy(coords0, q, dxy0, st.s));
  NNGP_TRY(dX.alloc(sizeof(double) * (size_t)n * p));
  NNGP_TRY(dX0.alloc(sizeof(double) * (size_t)q * p));
  NNGP_TRY(dnn0.alloc(sizeof(int) * (size_t)q * m));
  NNGP_TRY(dbeta.alloc(sizeof(double) * (size_t)p * nSamples));
  NNGP_TRY(dtheta.alloc(sizeof(double) * (size_t)nTheta * nSamples));
  NNGP_TRY(dw.alloc(sizeof(double) * (size_t)n * nSamples));
  NNGP_TRY(dz.alloc(sizeof(double) * 2 * qs));
  NNGP_TRY(dy0.alloc(sizeof(double) * qs));
  NNGP_TRY(dw0.alloc(sizeof(double) * qs));
  NNGP_TRY(derr.alloc(sizeof(int)));
  NNGP_HIP(hipMemcpyAsync(dX.p, X, dX.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dX0.p, X0, dX0.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dnn0.p, nnIndx0, dnn0.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dbeta.p, betaSamples, dbeta.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dtheta.p, thetaSamples, dtheta.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dw.p, wSamples, dw.bytes, hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dz.p, z, dz.bytes, hipMemcpyHostToDevice, st.s));
  std::vector<CovArgs> hc;
  const bool generic = covModel == 2 || m > 16;
  if (generic) {
    hc.resize(nSamples);
    for (int s = 0; s < nSamples; s++) {
      const double *th = thetaSamples + (size_t)s * nTheta;
      hc[s] = make_cov(covModel, th[0], th[1], th[2], covModel == 2 ? th[3] : 0.0);
    }
    NNGP_TRY(dcov.alloc(sizeof(CovArgs) * hc.size()));
    NNGP_HIP(hipMemcpyAsync(dcov.p, hc.data(), dcov.bytes, hipMemcpyHostToDevice, st.s));
  }
  if (fill_int_launch(derr.as<int>(), INT_MAX, st.s))
    return set_err(NNGP_ERR_HIP, "flag reset kernel failed");
  PredArgs a;
  a.xy = dxy.as<double2>();
  a.X = dX.as<double>();
  a.y = nullptr;
  a.n = n;
  a.p = p;
  a.m = m;
  a.xy0 = dxy0.as<double2>();
  a.X0 = dX0.as<double>();
  a.q = q;
  a.nn0 = dnn0.as<int>();
  a.beta = dbeta.as<double>();
  a.theta = dtheta.as<double>();
  a.S = nSamples;
  a.nTheta = nTheta;
  a.cov = covModel;
  a.z = dz.as<double>();
  a.y0 = dy0.as<double>();
  a.err = derr.as<int>();
  a.unit = make_cov(covModel, 1.0, 0.0, 1.0, 0.0);
  a.cov_s = generic ? dcov.as<CovArgs>() : nullptr;
  a.wS = dw.as<double>();
  a.z2 = dz.as<double>() + qs;
  a.w0out = dw0.as<double>();
  const int chunk = 65535;
  for (int s0 = 0; s0 < q; s0 += chunk) {
    a.site0 = s0;
    a.nsite = std::min(chunk, q - s0);
    if (pred_launch(a, st.s)) return set_err(NNGP_ERR_HIP, "kriging launch failed");
  }
  int bad = INT_MAX;
  NNGP_HIP(hipMemcpyAsync(pY0, dy0.p, dy0.bytes, hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(pW0, dw0.p, dw0.bytes, hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(&bad, derr.p, sizeof(int), hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipStreamSynchronize(st.s));
  if (bad != INT_MAX)
    return set_err(NNGP_ERR_NOT_PD,
                   "c++ Rf_error: dpotrf failed\n(kriging system of new site %d"
                   " is not positive definite)", bad);
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// device-resident model
// ---------------------------------------------------------------------------
struct nngp_model {
  Model md;
};

int nngp_model_create(nngp_model **out, int n, int m, int p,
                      const double *coords, const int *nnIndx,
                      const int *nnIndxLU, const double *X, const double *y,
                      int covModel, const nngp_opts *opts) {
  clear_err();
  if (!out) return set_err(NNGP_ERR_ARG, "out is NULL");
  *out = nullptr;
  NNGP_TRY(check_cov(covModel));
  NNGP_TRY(check_nm(n, m));
  if (p < 0 || p > kMaxP || (p > 0 && (!X || !y)))
    return set_err(NNGP_ERR_ARG, "p must be in 0..%d with X and y given", kMaxP);
  if (!coords) return set_err(NNGP_ERR_ARG, "coords is NULL");
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  NNGP_TRY(use_device(opts));
  nngp_model *h = new (std::nothrow) nngp_model;
  if (!h) return set_err(NNGP_ERR_NOMEM, "host allocation failed");
  const int st = h->md.create(n, m, p, coords, nnIndx, nnIndxLU,
                              p > 0 ? X : nullptr, p > 0 ? y : nullptr,
                              covModel, opts);
  if (st != NNGP_OK) {
    delete h;
    return st;
  }
  *out = h;
  return NNGP_OK;
}

void nngp_model_destroy(nngp_model *model) { delete model; }

int nngp_model_update(nngp_model *model, const double *theta, int slot,
                      const double *beta, int want_xprod,
                      nngp_pass_out *out) {
  clear_err();
  if (!model || !theta || slot < 0 || slot > 1)
    return set_err(NNGP_ERR_ARG, "bad model handle, theta or slot");
  Model &md = model->md;
  const CovArgs c = make_cov(md.cov, theta[0], theta[1], theta[2],
                             md.cov == 2 ? theta[3] : 0.0);
  PassOut po;
  NNGP_TRY(md.pass(c, slot, true, &po));
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->logDet = po.logDet;
    out->qResid = (md.nz && beta) ? quad_resid(po, beta, md.p) : 0.0;
    if (want_xprod && md.nz) {
      const int p = md.p;
      for (int i = 0; i < p; i++) {
        out->XtViy[i] = gram_at(po.gram, p, i);
        for (int j = 0; j < p; j++)
          out->XtViX[j * p + i] = i >= j ? gram_at(po.gram, i, j) : 0.0;
      }
      out->ytViy = gram_at(po.gram, p, p);
    }
  }
  return NNGP_OK;
}

int nngp_model_qresid(nngp_model *model, int slot, const double *beta,
                      double *q) {
  clear_err();
  if (!model || !beta || !q || slot < 0 || slot > 1)
    return set_err(NNGP_ERR_ARG, "bad arguments");
  Model &md = model->md;
  if (!md.nz) return set_err(NNGP_ERR_ARG, "model was created without X, y");
  *q = quad_resid(md.slot_out[slot], beta, md.p);
  return NNGP_OK;
}

int nngp_model_rNNGP(nngp_model *model, const double *sigmaSqIG,
                     const double *tauSqIG, const double *phiUnif,
                     const double *nuUnif, const double *betaStarting,
                     double sigmaSqStarting, double tauSqStarting,
                     double phiStarting, double nuStarting,
                     double sigmaSqTuning, double tauSqTuning,
                     double phiTuning, double nuTuning, int nSamples,
                     int verbose, int nReport, const nngp_rng *rng,
                     const nngp_opts *opts, double *pBetaSamples,
                     double *pThetaSamples, int *nAccept) {
  clear_err();
  if (!model) return set_err(NNGP_ERR_ARG, "model is NULL");
  Model &md = model->md;
  if (md.nz == 0 || md.p < 1)
    return set_err(NNGP_ERR_ARG, "model was created without X, y");
  if (!sigmaSqIG || !tauSqIG || !phiUnif || !betaStarting || !pBetaSamples ||
      !pThetaSamples || nSamples < 0 || (md.cov == 2 && !nuUnif))
    return set_err(NNGP_ERR_ARG, "NULL buffer or negative nSamples");
  NNGP_TRY(use_device(opts));
  if (verbose)
    print_model(md.n, md.p, md.m, md.cov, nSamples, sigmaSqIG, tauSqIG,
                phiUnif, nuUnif, opts ? opts->device : 0);
  Rng R(rng, opts ? opts->seed : 1ull);
  McmcSpec S;
  std::memset(&S, 0, sizeof S);
  S.sigmaSqIG = sigmaSqIG;
  S.tauSqIG = tauSqIG;
  S.phiUnif = phiUnif;
  S.nuUnif = nuUnif;
  S.betaStarting = betaStarting;
  S.start[0] = sigmaSqStarting;
  S.start[1] = tauSqStarting;
  S.start[2] = phiStarting;
  S.start[3] = nuStarting;
  S.tuning[0] = sigmaSqTuning;
  S.tuning[1] = tauSqTuning;
  S.tuning[2] = phiTuning;
  S.tuning[3] = nuTuning;
  S.nSamples = nSamples;
  S.verbose = verbose;
  S.nReport = nReport;
  S.pBetaSamples = pBetaSamples;
  S.pThetaSamples = pThetaSamples;
  S.nAccept = nAccept;
  return mcmc_loop(md, S, R, opts);
}

int nngp_model_profile(nngp_model *model, int enable) {
  clear_err();
  if (!model) return set_err(NNGP_ERR_ARG, "model is NULL");
  Model &md = model->md;
  md.prof = enable != 0;
  md.prof_every = enable > 1 ? enable : 1;
  md.prof_seq = 0;
  md.prof_sample = false;
  md.nev = 0;
  md.tms[0] = md.tms[1] = 0.0;
  md.tcount[0] = md.tcount[1] = 0;
  md.loop_ms = 0.0;
  md.loop_iters = 0;
  return NNGP_OK;
}

int nngp_model_loop_timing(nngp_model *model, double *t) {
  clear_err();
  if (!model || !t) return set_err(NNGP_ERR_ARG, "bad arguments");
  t[0] = model->md.loop_ms;
  t[1] = (double)model->md.loop_iters;
  return NNGP_OK;
}

int nngp_model_timing(nngp_model *model, double *t) {
  clear_err();
  if (!model || !t) return set_err(NNGP_ERR_ARG, "bad arguments");
  Model &md = model->md;
  NNGP_HIP(hipStreamSynchronize(md.st.s));
  NNGP_TRY(md.collect());
  t[0] = md.tms[0];
  t[1] = (double)md.tcount[0];
  t[2] = md.tms[1];
  t[3] = (double)md.tcount[1];
  return NNGP_OK;
}

int nngp_model_autotune(nngp_model *model, const double *theta,
                        nngp_tune_report *report) {
  clear_err();
  if (!model || !theta) return set_err(NNGP_ERR_ARG, "bad arguments");
  Model &md = model->md;
  const CovArgs c = make_cov(md.cov, theta[0], theta[1], theta[2],
                             md.cov == 2 ? theta[3] : 0.0);
  BfTuneReport rep;
  NNGP_TRY(md.autotune(c, &rep));
  if (report) {
    report->n = rep.n;
    for (int k = 0; k < NNGP_TUNE_MAX; k++) {
      report->family[k] = k < rep.n ? rep.family[k] : 0;
      report->threads[k] = k < rep.n ? rep.threads[k] : 0;
      report->mode[k] = k < rep.n ? rep.mode[k] : 0;
      report->ms[k] = k < rep.n ? rep.ms[k] : 0.0;
    }
    report->best_family = rep.best.family;
    report->best_threads = rep.best.threads;
    report->best_mode = rep.best.mode;
  }
  return NNGP_OK;
}

int nngp_model_set_shape(nngp_model *model, int family, int threads,
                         int mode) {
  clear_err();
  if (!model) return set_err(NNGP_ERR_ARG, "bad arguments");
  return model->md.pin_shape(family, threads, mode);
}

int nngp_model_get_bf(nngp_model *model, int slot, double *B, double *F) {
  clear_err();
  if (!model || slot < 0 || slot > 1) return set_err(NNGP_ERR_ARG, "bad arguments");
  Model &md = model->md;
  if (B) {
    if (!md.B[slot].p) return set_err(NNGP_ERR_ARG, "slot %d holds no B yet", slot);
    NNGP_HIP(hipMemcpy(B, md.B[slot].p, sizeof(double) * (size_t)md.nIndx,
                       hipMemcpyDeviceToHost));
  }
  if (F)
    NNGP_HIP(hipMemcpy(F, md.F[slot].p, sizeof(double) * (size_t)md.n,
                       hipMemcpyDeviceToHost));
  return NNGP_OK;
}

}  // extern "C"
