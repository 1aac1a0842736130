/*
 * nngp_oracle.h — CPU oracle for the NNGP hot path of spNNGP 1.0.1.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference's
 * algorithms, used as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product (spnngp_amd, libnngp_hip.so)
 * never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Layouts are the
 * reference's: column-major FP64, int32 indices, flat nnIndx/nnIndxLU.
 *
 * Parity pinning (tests/test_oracle_*.py):
 *   or_getNNIndx      exhaustive offsets/counts vs the closed form
 *   or_spCor          closed forms (exp, gaussian, spherical branches,
 *                     matern nu = 0.5/1.5/2.5), general nu vs scipy kv
 *   or_bessel_k       scipy.special.kv, half-integer closed forms
 *   or_mkNNIndx(CB)   brute force in numpy, scipy cKDTree on prefixes,
 *                     CB == brute on tie-free data
 *   or_updateBF       residual ||C B - c||, scipy potrf/potri/symv sequence,
 *                     exact-GP identity (m >= n-1: NNGP likelihood equals the
 *                     dense Gaussian log-density)
 *   or_Q              dense (I-B)^T F^-1 (I-B) quadratic form
 *   or_rNNGP          RNG draw order by hand-rolled single steps, posterior
 *                     recovery; dense-GP likelihood at m >= n-1
 *   or_rNNGPPredict   dense GP kriging when m = n (z = 0 isolates the mean)
 *   or_nngp_draw,     at m >= n-1 the draw is v = W z with W W^T equal to the
 *   or_rNNGPReplicated  dense covariance s2 R + t2 I (W = (I-B)^-1 F^1/2)
 *   or_cNNGP          dense GLS when m >= n-1
 *   or_philox*        Random123 known-answer vectors; moments and a KS test
 *   or_mkUIndx        the three search types against each other and a
 *                     numpy transpose; every (child, slot) names the site
 *   or_rgamma         Kolmogorov-Smirnov against scipy's gamma CDF over
 *                     shapes in every branch of algorithm GD
 *   or_sNNGPPredict   dense GP conditional of w at m = n (mean from z = 0,
 *                     sd from z = 1) and the tau.sq noise term
 *   or_sNNGP          the w update equals the full conditional of w_i from
 *                     the dense precision (I-B)^T F^-1 (I-B) + I / tau.sq;
 *                     a chain under a permuted visit order equals the
 *                     permuted-scan sampler; posterior recovery
 * Parity unpinned (no R in this image): R's RNG stream, the exact bits of
 * Rmath's bessel_k_ex / gammafn, R's LAPACK rounding order, and the order
 * rsort_with_index gives to exactly tied distances.
 */
#ifndef NNGP_ORACLE_H
#define NNGP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_rng {
  void *state;
  double (*norm_rand)(void *state);
  double (*unif_rand)(void *state);
} or_rng;

/* status codes (same numbers as the product's) */
#define OR_OK 0
#define OR_ERR_ARG 1
#define OR_ERR_NOT_PD 2
#define OR_ERR_COV_MODEL 3

const char *or_last_error(void);

/* src/util.cpp:55-70 */
void or_getNNIndx(int i, int m, int *iNNIndx, int *iNN);

/* src/util.cpp:51-53 */
double or_dist2(double a1, double a2, double b1, double b2);

/* K_nu(x), x > 0 (Rmath bessel_k_ex with expo = 1; Temme series / Steed
 * continued fraction + upward recurrence). */
double or_bessel_k(double x, double nu);

/* src/util.cpp:208-245.  Returns NaN for a covModel outside 0..3. */
double or_spCor(double D, double phi, double nu, int covModel);

/* R's rsort_with_index (shell sort, gaps 1, 4, 13, ...; used at
 * src/nn.cpp:104,176,197,243). */
void or_rsort_with_index(double *x, int *indx, int n);

/* src/nn.cpp:63-111 (brute force). */
int or_mkNNIndx(int n, int m, const double *coords, int *nnIndx,
                double *nnDist, int *nnIndxLU, int nThreads);

/* src/nn.cpp:128-257 (code-book / u = x + y walk). */
int or_mkNNIndxCB(int n, int m, const double *coords, int *nnIndx,
                  double *nnDist, int *nnIndxLU, int nThreads);

/* Exact k nearest of q query sites among n sites (RANN nn2 semantics,
 * R/generics.R:642); ties by ascending index.  nnIndx0 q x k col-major. */
int or_nn2(int n, const double *coords, int q, const double *coords0, int k,
           int *nnIndx0, double *nnDist0);

/* LAPACK/BLAS-style unblocked routines on the lower triangle,
 * column-major, leading dimension n (dpotf2, dtrti2 + dlauu2, dsymv). */
int or_potrf_L(int n, double *A);
int or_potri_L(int n, double *A);
void or_symv_L(int n, const double *A, const double *x, double *y);

/* src/rNNGP.cpp:46-101.  theta = {sigma.sq, tau.sq, phi, nu}. */
int or_updateBF(double *B, double *F, const double *coords,
                const int *nnIndx, const int *nnIndxLU, int n, int m,
                const double *theta, int covModel, int nThreads,
                double *logDet);

/* src/cNNGP.cpp:24-69. */
int or_updateConjBF(double *B, double *F, const double *coords,
                    const int *nnIndx, const int *nnIndxLU, int n, int m,
                    double phi, double alpha, double nu, int covModel,
                    int nThreads);

/* src/util.cpp:261-280, summed serially in site order. */
double or_Q(const double *B, const double *F, const double *u,
            const double *v, int n, const int *nnIndx, const int *nnIndxLU);

/* src/rNNGP.cpp:105-420 (+ updateRep, src/rNNGP.cpp:24-42). */
int or_rNNGP(const double *y, const double *X, int p, int n, int m,
             const double *coords, int covModel, const int *nnIndx,
             const int *nnIndxLU, const double *sigmaSqIG,
             const double *tauSqIG, const double *phiUnif,
             const double *nuUnif, const double *betaStarting,
             double sigmaSqStarting, double tauSqStarting,
             double phiStarting, double nuStarting, double sigmaSqTuning,
             double tauSqTuning, double phiTuning, double nuTuning,
             int nSamples, int nThreads, int nRep, const int *repIndx,
             const or_rng *rng, double *pBetaSamples, double *pThetaSamples,
             double *yRepSamples, int *nAccept);

/* src/rNNGPPredict.cpp:25-222 with z supplied site-major (z[i*S+s]). */
int or_rNNGPPredict(const double *X, const double *y, const double *coords,
                    int n, int p, int m, const double *X0,
                    const double *coords0, int q, const int *nnIndx0,
                    const double *betaSamples, const double *thetaSamples,
                    int nSamples, int covModel, int nThreads, const double *z,
                    double *pY0);

/* src/cNNGP.cpp:73-358; ab is written as 2 x g. */
int or_cNNGP(const double *y, const double *X, int p, int n,
             const double *coords, const double *thetaAlpha,
             const int *nnIndx, const int *nnIndxLU, const double *X0,
             const double *coords0, int n0, const int *nnIndx0, int g, int m,
             const double *sigmaSqIG, int covModel, int nThreads,
             double *betaHat, double *ab, double *bBInv, double *bb,
             double *y0Hat, double *y0HatVar);

/* Sequential NNGP draw w_i = B_i . w_N(i) + sqrt(F_i) z_i with z supplied
 * (src/rNNGP.cpp:24-42, src/rNNGPReplicated.cpp:183-199). */
void or_nngp_draw(const double *B, const double *F, const int *nnIndx,
                  const int *nnIndxLU, int n, const double *z, double *w);

/* src/rNNGPReplicated.cpp:78-230: for each sample s, B and F of
 * theta[s*nTheta..] (or_updateBF, tau.sq on the diagonal), the draw of
 * or_nngp_draw with n normals from rng (site order), then
 * yRep[s*n + i] = v_i + sum_j X[i + j*n] beta[s + j*nSamples]. */
int or_rNNGPReplicated(const double *X, int p, int n, int m,
                       const double *coords, int covModel, const int *nnIndx,
                       const int *nnIndxLU, const double *beta,
                       const double *theta, int nSamples, int nThreads,
                       const or_rng *rng, double *yRep);

/* Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11) in place on the
 * counter c[4] under key (k0, k1), and the standard normal the kriging
 * server draws for site i, sample s of a request with key `seed`
 * (counter (i, s, 0, 0), two 53-bit uniforms, Box-Muller cosine branch).
 * Not a reference function: the reference draws z from R's RNG
 * (src/rNNGPPredict.cpp:164).  Pinned by Random123's known-answer
 * vectors and by normality of its output. */
void or_philox4x32_10(unsigned int c[4], unsigned int k0, unsigned int k1);
double or_philox_normal(unsigned long long seed, unsigned int i,
                        unsigned int s);

/* src/nn.cpp:25-56 (mkUIndx): for each site i the sites that hold i as a
 * neighbor, uIndx[uIndxLU[i] .. + uIndxLU[n+i]), and i's slot in each of
 * their neighbor lists, uiIndx.  searchType 0 / 1: the scans of
 * src/util.cpp:72-112 (children ascending / descending); otherwise the
 * CSR -> CSC transpose of src/util.cpp:115-185 (ascending). */
int or_mkUIndx(int n, int m, const int *nnIndx, int *uIndx, int *uIndxLU,
               int *uiIndx, const int *nnIndxLU, int searchType);

/* Gamma(a, scale) deviate, a >= 1, by Ahrens & Dieter's algorithm GD (the
 * method of R's rgamma for a >= 1; exponential deviates by their algorithm
 * SA, as R's exp_rand).  NaN for a < 1. */
double or_rgamma(const or_rng *r, double a, double scale);

/* src/sNNGP.cpp:90-482, the latent-model sampler.  visit == NULL sweeps w
 * in site order (the reference); otherwise in the order visit[0..n-1]
 * (site i always takes the i-th of the n normals drawn before the sweep).
 * pBetaSamples p x S, pThetaSamples nTheta x S, pWSamples n x S. */
int or_sNNGP(const double *y, const double *X, int p, int n, int m,
             const double *coords, int covModel, const int *nnIndx,
             const int *nnIndxLU, const int *uIndx, const int *uIndxLU,
             const int *uiIndx, const double *sigmaSqIG,
             const double *tauSqIG, const double *phiUnif,
             const double *nuUnif, const double *betaStarting,
             double sigmaSqStarting, double tauSqStarting,
             double phiStarting, double nuStarting, double phiTuning,
             double nuTuning, int nSamples, int nThreads, const int *visit,
             const or_rng *rng, double *pBetaSamples, double *pThetaSamples,
             double *pWSamples, int *nAccept);

/* src/sNNGPPredict.cpp:25-241 (Gaussian family, single-thread normal
 * order): w0[s*q+i] = sqrt(s2 - u.c) z[i*S+s] + u . w_s(N),
 * y0[s*q+i] = sqrt(t2) z[q*S + i*S+s] + x0_i . beta_s + w0[s*q+i],
 * u = (s2 R(N,N))^-1 s2 R(N,i).  w n x S, beta p x S, theta nTheta x S. */
int or_sNNGPPredict(const double *X0, const double *coords, int n, int p,
                    int m, const double *coords0, int q, const int *nnIndx0,
                    const double *beta, const double *theta, const double *w,
                    int nSamples, int covModel, const double *z, double *y0,
                    double *w0);

#ifdef __cplusplus
}
#endif
#endif
