/*
 * nngp.h — C ABI of the MI355X (gfx950) NNGP library, libnngp_hip.so.
 *
 * Drop-in boundary for spNNGP 1.0.1's NNGP hot path.  Every entry point takes
 * plain pointers and sizes (no R, no torch types).  Matrices keep the
 * reference's column-major FP64 layouts and int32 index layouts, so an R
 * `.Call` shim is pure pointer pass-through.
 *
 * Conventions shared by all entry points
 *   - n sites, m = n.neighbors, p covariates, m_i = min(i, m).
 *   - coords: n x 2 column-major (x in [0,n), y in [n,2n)), already in NNGP
 *     order for the fitting routines (R/spNNGP.R:144).
 *   - nnIndx / nnIndxLU: flat neighbor layout of src/util.cpp:55-70; site i's
 *     neighbors are nnIndx[nnIndxLU[i] .. nnIndxLU[i]+nnIndxLU[n+i]).
 *     nIndx = m(m+1)/2 + (n-m-1)m  (src/rNNGP.cpp:218).
 *   - covModel: 0 exponential, 1 spherical, 2 matern, 3 gaussian
 *     (src/util.cpp:192-206).
 *   - theta for the response model: {sigma.sq, tau.sq, phi[, nu]}
 *     (src/rNNGP.cpp:182-190).
 *   - Return value: NNGP_OK (0) or an NNGP_ERR_* code; nngp_last_error()
 *     then holds a message.  Where the reference raises Rf_error the message
 *     starts with the reference's text, e.g. "c++ Rf_error: dpotrf failed".
 *   - Caller owns every host buffer.  The library owns device memory; every
 *     call allocates what it needs and frees it before returning, except the
 *     nngp_model_* handle, which keeps the fitted data resident in HBM.
 *   - Calls on one nngp_model are not thread-safe; distinct handles may be
 *     used from distinct host threads.
 */
#ifndef NNGP_H
#define NNGP_H

#ifdef __cplusplus
extern "C" {
#endif

enum nngp_status {
  NNGP_OK = 0,
  NNGP_ERR_ARG = 1,         /* bad size, null pointer, malformed nnIndxLU   */
  NNGP_ERR_NOT_PD = 2,      /* a neighbor covariance failed Cholesky        */
  NNGP_ERR_COV_MODEL = 3,   /* covModel outside 0..3                        */
  NNGP_ERR_HIP = 4,         /* HIP runtime error (no device, launch, copy)  */
  NNGP_ERR_NOMEM = 5,       /* device or host allocation failed             */
  NNGP_ERR_INTERRUPT = 6,   /* the caller's interrupt callback said stop    */
  NNGP_ERR_UNSUPPORTED = 7, /* configuration the library does not handle    */
  NNGP_ERR_TIMEOUT = 8      /* nngp_server_wait: nothing finished in time   */
};

/* Order of the latent sampler's w sweep (nngp_opts.sweep). */
enum nngp_sweep {
  NNGP_SWEEP_SITE_ORDER = 0, /* sites 0..n-1, the reference's Gibbs scan    */
  NNGP_SWEEP_COLOURED = 1    /* colour classes of the moral graph, one
                                parallel phase per colour                   */
};

enum nngp_precision {
  NNGP_FP64 = 0,            /* plain FP64: the reference's arithmetic       */
  NNGP_MIXED = 1            /* FP32 factor/solve + one FP64 refinement step */
};

/* Host random number source.  The library never draws random numbers on
 * the device for the MCMC; it calls these between kernels, in the order of
 * src/rNNGP.cpp (beta normals, proposal normals, one uniform).  An R shim
 * passes wrappers of norm_rand()/unif_rand() between GetRNGstate() and
 * PutRNGstate().  NULL selects the library's own generator, seeded by
 * nngp_opts.seed. */
typedef struct nngp_rng {
  void *state;
  double (*norm_rand)(void *state);
  double (*unif_rand)(void *state);
} nngp_rng;

/* R's default generator (RNGkind "Mersenne-Twister", normal.kind
 * "Inversion") as an nngp_rng, seeded the way set.seed(seed) seeds it, for
 * callers without R that want R's stream (no counterpart in the reference,
 * which calls R's own generator).  nngp_rng_r_destroy frees the state. */
int nngp_rng_r_create(unsigned int seed, nngp_rng *out);
void nngp_rng_r_destroy(nngp_rng *rng);

/* Per-call options.  NULL means all defaults (FP64, device 0, seed 1). */
typedef struct nngp_opts {
  int precision;            /* enum nngp_precision                          */
  int device;               /* HIP device ordinal                           */
  unsigned long long seed;  /* seed of the built-in generator (rng == NULL) */
  int (*interrupt)(void *); /* polled once per MCMC iteration / grid point;
                               nonzero stops with NNGP_ERR_INTERRUPT        */
  void *interrupt_arg;
  int shape[3];             /* updateBF launch shape (family, threads, mode
                               as in nngp_tune_report) pinned for a model or
                               a grid context; threads == 0: the default,
                               or the autotuner's choice                    */
  int sweep;                /* enum nngp_sweep, for nngp_sNNGP              */
  int search;               /* a grid context's per-fold neighbor search:
                               0 code book (mkNNIndxCB), 1 brute (mkNNIndx),
                               R's search.type (R/spConjNNGP.R:301-305)    */
  int reserved[3];
} nngp_opts;

/* Message for the last nonzero status returned on this host thread. */
const char *nngp_last_error(void);

/* Library version, e.g. 10000 for 1.0.0. */
int nngp_version(void);

/* Number of HIP devices visible; 0 when none (never an error). */
int nngp_device_count(void);

/* ------------------------------------------------------------------------
 * Neighbor search
 * --------------------------------------------------------------------- */

/* mkNNIndxCB — replaces src/nn.cpp:214-257 (.Call by name from R/nn.R:71).
 * Exact m nearest neighbors of every site among its NNGP predecessors
 * (sites with a smaller index), ascending by squared distance
 * fl(fl(dx*dx)+fl(dy*dy)); nnDist holds the Euclidean distances.
 * Fills caller-allocated nnIndx[nIndx], nnDist[nIndx], nnIndxLU[2n] in
 * place.  nThreads is accepted for signature parity and ignored. */
int nngp_mkNNIndxCB(int n, int m, const double *coords, int *nnIndx,
                    double *nnDist, int *nnIndxLU, int nThreads);

/* mkNNIndx — replaces the brute-force twin src/nn.cpp:63-111: the same
 * neighbor sets as nngp_mkNNIndxCB where distances do not tie.  Where they
 * do, each search gives the reference's own result: the code-book walk's
 * candidate order and R's shell sort (rsort_with_index) decide the order
 * of tied neighbors and which of them is kept at the m-th place for
 * mkNNIndxCB, the ascending scan and the same sort for mkNNIndx (which
 * compares Euclidean, not squared, distances). */
int nngp_mkNNIndx(int n, int m, const double *coords, int *nnIndx,
                  double *nnDist, int *nnIndxLU, int nThreads);

/* mkUIndx — replaces src/nn.cpp:25-56 (.Call at R/nn.R:19; the scans
 * mkUIndx0/1/2 of src/util.cpp:72-185 and the uiIndx fill of
 * src/nn.cpp:46-53).  For each site i, the sites holding i as a neighbor
 * are uIndx[uIndxLU[i] .. uIndxLU[i]+uIndxLU[n+i]), and uiIndx at the same
 * position is i's slot in that site's neighbor list.  searchType 0 and 2
 * list them in ascending order, 1 in descending order (the reference's
 * "opposite sorting among the children").  Caller-allocated: uIndx[nIndx],
 * uIndxLU[2n], uiIndx[nIndx].  Runs on the device (a stable radix sort of
 * the neighbor entries). */
int nngp_mkUIndx(int n, int m, const int *nnIndx, int *uIndx, int *uIndxLU,
                 int *uiIndx, const int *nnIndxLU, int searchType);

/* nn2 replacement (RANN, called at R/generics.R:642, R/spConjNNGP.R:158,311):
 * exact k nearest fitted sites of each of q query sites, no ordering
 * constraint, ascending by squared distance fl(fl(dx*dx)+fl(dy*dy)) and,
 * among equal distances, by ascending fitted row (RANN's own order among
 * exact ties depends on its kd-tree and is not reproduced).  nnIndx0 is
 * q x k column-major, 0-based (the R code's `nn2(...)$nn.idx - 1`);
 * nnDist0 (may be NULL) likewise. */
int nngp_nn2(int n, const double *coords, int q, const double *coords0,
             int k, int *nnIndx0, double *nnDist0);

/* ------------------------------------------------------------------------
 * Numerical kernels (src/rNNGP.cpp:46-101, src/cNNGP.cpp:24-69,
 * src/util.cpp:261-280), host buffers in and out.
 * --------------------------------------------------------------------- */

/* updateBF — replaces src/rNNGP.cpp:46-101.  theta = {sigma.sq, tau.sq,
 * phi[, nu]}.  Writes B[nIndx], F[n]; *logDet = sum_i log F_i.  B is not
 * written for site 0 (it has no neighbors). */
int nngp_updateBF(double *B, double *F, const double *coords,
                  const int *nnIndx, const int *nnIndxLU, int n, int m,
                  const double *theta, int covModel, double *logDet,
                  const nngp_opts *opts);

/* updateConjBF — replaces src/cNNGP.cpp:24-69 (correlation scale:
 * C = R + alpha I, F = 1 + alpha - B.c). */
int nngp_updateConjBF(double *B, double *F, const double *coords,
                      const int *nnIndx, const int *nnIndxLU, int n, int m,
                      double phi, double alpha, double nu, int covModel,
                      const nngp_opts *opts);

/* Q — replaces src/util.cpp:261-280:
 * sum_i (u_i - B_i.u_N(i)) (v_i - B_i.v_N(i)) / F_i. */
int nngp_Q(const double *B, const double *F, const double *u,
           const double *v, int n, int m, const int *nnIndx,
           const int *nnIndxLU, double *q);

/* Q on device-resident buffers, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * B, F, u, v, nnIndx, nnIndxLU are device pointers in nngp_Q's layouts
 * (not validated: the caller guarantees them); *q, also on the device,
 * receives the sum.  The kernels' partial sums live in a library-owned
 * device array of 1024 doubles (8 KiB), held by every GPU the library runs
 * on until the process ends; the reduction grid never has more than 1024
 * blocks, so every n whose layout fits int32 (nIndx < 2^31) is accepted
 * and larger ones are refused.  The work is enqueued on `stream` (a
 * hipStream_t of the current device; NULL: the null stream) and not waited
 * for.  Calls on one GPU share the array: calls that overlap on one GPU
 * race, so order them (one stream, or events). */
int nngp_Q_noalloc(const double *B, const double *F, const double *u,
                   const double *v, int n, int m, const int *nnIndx,
                   const int *nnIndxLU, double *q, void *stream);

/* Sequential NNGP draw w_i = B_i . w_N(i) + sqrt(F_i) z_i over i = 0..n-1
 * — replaces updateRep, src/rNNGP.cpp:24-42 (also
 * src/rNNGPReplicated.cpp:183-199), with the n standard normals z supplied
 * by the caller.  Host code (each w_i depends on earlier ones). */
int nngp_prior_draw(const double *B, const double *F, const int *nnIndx,
                    const int *nnIndxLU, int n, const double *z, double *w);

/* sNNGPPredict — replaces src/sNNGPPredict.cpp:25-241 (.Call at
 * R/generics.R:671-672), predictions from a latent fit.  For every new
 * site i and posterior sample s, with the m fitted neighbors N of i:
 * C = sigma.sq R(N,N) (no nugget), c = sigma.sq R(N,i), u = C^-1 c,
 *   pW0[s*q + i] = sqrt(sigma.sq - u.c) z[i*S + s] + u . w_s(N),
 *   pY0[s*q + i] = sqrt(tau.sq) z[q*S + i*S + s] + x0_i . beta_s + pW0[s*q+i]
 * with w_s = wSamples[s*n ..] (n x nSamples).  betaSamples p x S,
 * thetaSamples nTheta x S ({sigma.sq, tau.sq, phi[, nu]}), nnIndx0 q x m.
 * z: the 2 q S normals in the reference's order (all of w0's, then all of
 * y0's), or NULL to draw them from rng (or the built-in generator).
 * family 1 (Gaussian) only; y is accepted for signature parity.  A 1-thread
 * reference run assigns the normals in this order; with more threads the
 * reference's assignment depends on scheduling (src/sNNGPPredict.cpp:185). */
int nngp_sNNGPPredict(const double *X, const double *y, const double *coords,
                      int n, int p, int m, const double *X0,
                      const double *coords0, int q, const int *nnIndx0,
                      const double *betaSamples, const double *thetaSamples,
                      const double *wSamples, int nSamples, int family,
                      int covModel, int nThreads, int verbose, int nReport,
                      const double *z, const nngp_rng *rng,
                      const nngp_opts *opts, double *pY0, double *pW0);

/* rNNGPReplicated — replaces src/rNNGPReplicated.cpp:78-230 (.Call at
 * R/generics.R:147, fitted() of a conjugate fit).  For each sample s: B and
 * F of theta_s (C = sigma.sq R + tau.sq I), the draw
 * v_i = B_i . v_N(i) + sqrt(F_i) z_{s,i} over i = 0..n-1, and
 * yRep[s*n + i] = v_i + sum_j X[i + j*n] beta[s + j*nSamples].
 * X n x p, coords n x 2 (NNGP order), beta nSamples x p, theta nTheta x
 * nSamples, all column-major; yRep n x nSamples.  The normals come from
 * rng in the reference's order (all n of sample s, then sample s + 1);
 * with rng == NULL they are Philox normals keyed by (opts->seed, site,
 * sample), as nngp_server_submit draws them.  Runs on the device: the
 * draw follows the neighbor DAG level by level, one sample per block, for
 * groups of up to 256 samples whose B, F and draw buffers (about
 * 8 (nnIndx entries + 4n) bytes a sample) take at most half the free device
 * memory (at least 16 GB). */
int nngp_rNNGPReplicated(const double *X, int p, int n, int m,
                         const double *coords, int covModel,
                         const int *nnIndx, const int *nnIndxLU,
                         const double *beta, const double *theta,
                         int nSamples, int verbose, int nReport,
                         const nngp_rng *rng, const nngp_opts *opts,
                         double *yRep);

/* ------------------------------------------------------------------------
 * .Call drivers (the argument lists follow the SEXP order of
 * src/spNNGP.h; outputs are caller-allocated)
 * --------------------------------------------------------------------- */

/* rNNGP — replaces src/rNNGP.cpp:105-420 (.Call at R/spNNGP.R:401-405).
 * Response-model Metropolis-within-Gibbs.  The 28 reference arguments in
 * order, then:
 *   rng             host random source (NULL: built-in, opts->seed)
 *   pBetaSamples    p x nSamples        (column s = iteration s)
 *   pThetaSamples   nTheta x nSamples   (nTheta = 3, or 4 for matern)
 *   yRepSamples     n x nRep, or NULL when nRep == 0
 *   nAccept         optional: number of accepted proposals
 * verbose prints the reference's progress lines on stdout. */
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
               double *pThetaSamples, double *yRepSamples, int *nAccept);

/* sNNGP — replaces src/sNNGP.cpp:90-482 (.Call at R/spNNGP.R:409-413).
 * Latent-model sampler: per iteration the w sweep (n normals, drawn in
 * site order), beta | w (p normals), tau.sq and sigma.sq from their
 * inverse-Gamma conditionals (one Gamma deviate each, by R's method), and a
 * Metropolis step on phi [and nu].  The 29 reference arguments in order
 * (uIndx, uIndxLU, uiIndx as nngp_mkUIndx writes them; the sigma.sq and
 * tau.sq tunings are accepted and unused, as in the reference), then:
 *   rng, opts       as for nngp_rNNGP; opts->sweep picks the w sweep order
 *                   (NNGP_SWEEP_SITE_ORDER reproduces the reference's
 *                   chain; NNGP_SWEEP_COLOURED is a colour-by-colour scan)
 *   pBetaSamples    p x nSamples
 *   pThetaSamples   nTheta x nSamples
 *   pWSamples       n x nSamples
 *   nAccept         optional: accepted phi [, nu] proposals
 * With rng == NULL the sweep's normals are Philox normals keyed by
 * (opts->seed, site, iteration) and drawn on the device; the other draws
 * use the built-in generator. */
int nngp_sNNGP(const double *y, const double *X, int p, int n, int m,
               const double *coords, int covModel, const int *nnIndx,
               const int *nnIndxLU, const int *uIndx, const int *uIndxLU,
               const int *uiIndx, const double *sigmaSqIG,
               const double *tauSqIG, const double *phiUnif,
               const double *nuUnif, const double *betaStarting,
               double sigmaSqStarting, double tauSqStarting,
               double phiStarting, double nuStarting, double sigmaSqTuning,
               double tauSqTuning, double phiTuning, double nuTuning,
               int nSamples, int nThreads, int verbose, int nReport,
               const nngp_rng *rng, const nngp_opts *opts,
               double *pBetaSamples, double *pThetaSamples,
               double *pWSamples, int *nAccept);

/* cNNGP — replaces src/cNNGP.cpp:73-358 (.Call at R/spConjNNGP.R:336,378).
 * thetaAlpha is nTheta x g ({phi, alpha[, nu]} per column).  Outputs:
 *   betaHat  p x g,  ab  2 x g (a, b per grid point),  bBInv  p*p x g,
 *   bb  p x g,  y0Hat  n0 x g,  y0HatVar  n0 x g (both unused if n0 == 0).
 * The reference stores ab at [k*2], [k*2+1] inside an nTheta x g matrix;
 * this entry point writes the same values into a 2 x g array. */
int nngp_cNNGP(const double *y, const double *X, int p, int n,
               const double *coords, const double *thetaAlpha,
               const int *nnIndx, const int *nnIndxLU, const double *X0,
               const double *coords0, int n0, const int *nnIndx0, int g,
               int m, const double *sigmaSqIG, int covModel, int nThreads,
               int verbose, const nngp_opts *opts, double *betaHat,
               double *ab, double *bBInv, double *bb, double *y0Hat,
               double *y0HatVar);

/* rNNGPPredict — replaces src/rNNGPPredict.cpp:25-222 (.Call at
 * R/generics.R:674).  betaSamples p x nSamples, thetaSamples
 * nTheta x nSamples, nnIndx0 q x m column-major.  z (q*nSamples standard
 * normals, site-major: z[i*nSamples+s], the reference's single-thread
 * order) may be NULL, in which case z is drawn from rng (or the built-in
 * generator) in that order.  Output pY0 is q x nSamples. */
int nngp_rNNGPPredict(const double *X, const double *y, const double *coords,
                      int n, int p, int m, const double *X0,
                      const double *coords0, int q, const int *nnIndx0,
                      const double *betaSamples, const double *thetaSamples,
                      int nSamples, int covModel, int nThreads, int verbose,
                      int nReport, const double *z, const nngp_rng *rng,
                      const nngp_opts *opts, double *pY0);

/* ------------------------------------------------------------------------
 * Device-resident model: coordinates, neighbor index, X and y stay in HBM
 * across MCMC iterations.  nngp_rNNGP is built on this; it is exported for
 * callers that run their own sampler.
 * --------------------------------------------------------------------- */

typedef struct nngp_model nngp_model;

/* Quadratic forms produced by one pass (all over the same B, F). */
typedef struct nngp_pass_out {
  double logDet;            /* sum_i log F_i                               */
  double qResid;            /* Q(r, r), r = X beta - y                     */
  double XtViX[64];         /* p x p column-major, lower triangle filled   */
  double XtViy[8];          /* p                                            */
  double ytViy;             /* Q(y, y)                                      */
} nngp_pass_out;

int nngp_model_create(nngp_model **out, int n, int m, int p,
                      const double *coords, const int *nnIndx,
                      const int *nnIndxLU, const double *X, const double *y,
                      int covModel, const nngp_opts *opts);
void nngp_model_destroy(nngp_model *model);

/* B, F of theta into slot (0 or 1) with logDet; when beta != NULL also
 * Q(r, r) for r = X beta - y; when want_xprod != 0 also XtViX, XtViy,
 * ytViy.  p <= 8. */
int nngp_model_update(nngp_model *model, const double *theta, int slot,
                      const double *beta, int want_xprod,
                      nngp_pass_out *out);

/* Q(r, r) for r = X beta - y over the B, F already in slot. */
int nngp_model_qresid(nngp_model *model, int slot, const double *beta,
                      double *q);

/* Copy B[nIndx], F[n] of slot back to the host (B only after a pass that
 * stored it: nngp_model_update always does). */
int nngp_model_get_bf(nngp_model *model, int slot, double *B, double *F);

/* The rNNGP sampler (src/rNNGP.cpp:192-388) on a model created with X and
 * y, so that the data is already resident when the chain starts.  Same
 * arguments and outputs as nngp_rNNGP without the data and without fit.rep
 * (nRep = 0). */
int nngp_model_rNNGP(nngp_model *model, const double *sigmaSqIG,
                     const double *tauSqIG, const double *phiUnif,
                     const double *nuUnif, const double *betaStarting,
                     double sigmaSqStarting, double tauSqStarting,
                     double phiStarting, double nuStarting,
                     double sigmaSqTuning, double tauSqTuning,
                     double phiTuning, double nuTuning, int nSamples,
                     int verbose, int nReport, const nngp_rng *rng,
                     const nngp_opts *opts, double *pBetaSamples,
                     double *pThetaSamples, int *nAccept);

/* Kernel timing on the model's stream (HIP events around each launch), for
 * the run-time breakdown the reference reports as run.time
 * (R/spNNGP.R:395,437).  enable != 0 starts (and zeroes) the counters;
 * enable = k > 1 times only every k-th updateBF launch and the sum launch
 * after it (each event is a packet on the stream: timing every launch slows
 * a chain of short passes by about 5%), 0 stops them.
 * nngp_model_timing fills t[0] = total updateBF kernel ms, t[1] = number of
 * timed updateBF launches, t[2] = total ms of the timed sum kernels, t[3] =
 * their launches. */
int nngp_model_profile(nngp_model *model, int enable);
int nngp_model_timing(nngp_model *model, double *t);

/* Host wall time of the MH samplers' iteration loops since timing was
 * enabled: t[0] = ms, t[1] = iterations.  Each loop is measured from its
 * first proposal to its last accept/reject decision, with the model's
 * stream drained at both ends; the call's set-up (the initial pass at the
 * starting theta) and the copy of the samples out are outside.  This is the
 * per-iteration rate of a running chain, which the reference reports only
 * as the whole call's run.time (R/spNNGP.R:395,437). */
int nngp_model_loop_timing(nngp_model *model, double *t);

/* Launch shapes of the updateBF pass timed by the autotuner (at most
 * NNGP_TUNE_MAX reported). */
#define NNGP_TUNE_MAX 12
typedef struct nngp_tune_report {
  int n;                          /* candidates timed                     */
  int family[NNGP_TUNE_MAX];      /* 0: one site per lane, 1: per wavefront,
                                     2: per 16-lane row, four sites a
                                     wavefront (m <= 32, FP64)            */
  int threads[NNGP_TUNE_MAX];     /* threads per block                    */
  int mode[NNGP_TUNE_MAX];        /* family 0, NNGP_MIXED: 1 = the
                                     covariance copy kept in LDS for the
                                     refinement residual, 3 = kept in
                                     registers (one wave per SIMD), 0 =
                                     re-evaluated; family 0, FP64:
                                     2 = the covariate rows of each site's
                                     neighbors staged in LDS during the
                                     factorization (same results bit for
                                     bit as 0)                            */
  double ms[NNGP_TUNE_MAX];       /* mean launch time                     */
  int best_family, best_threads, best_mode;
} nngp_tune_report;

/* Time every launch shape of this model's updateBF pass at theta (sites per
 * wavefront 64, 4 or 1; 64/128/256-thread blocks or 1/2/4-system LDS tiles)
 * and use the fastest for every later pass.  Opt-in: the shape fixes the
 * order of the per-block sums, so a chain is bit-reproducible for a fixed
 * shape; without this call the shape is fixed: the per-lane kernel in
 * 64-thread blocks with the covariate rows staged where it applies, else
 * for FP64 passes with m <= 32 the 16-lane-row kernel in 256-thread blocks,
 * else the wavefront kernel in its widest LDS tile. */
int nngp_model_autotune(nngp_model *model, const double *theta,
                        nngp_tune_report *report);

/* Pin the launch shape of this model's updateBF passes (family, threads,
 * mode as in nngp_tune_report), e.g. one chosen by an earlier autotune, so
 * that chains are bit-reproducible across processes.  NNGP_ERR_ARG when the
 * shape does not apply to the model (m, covModel, precision).  No
 * counterpart in the reference, which has a single CPU code path. */
int nngp_model_set_shape(nngp_model *model, int family, int threads,
                         int mode);

/* ------------------------------------------------------------------------
 * Kriging server: rNNGPPredict (src/rNNGPPredict.cpp:25-222) together with
 * its nn2 neighbor query (R/generics.R:642) as a resident service.  The
 * fitted response model (coordinates, X, y, posterior samples) and its
 * neighbor-search index stay in HBM; requests of any size are submitted
 * from any host thread and answered as they finish.
 *
 * Continuous batching: a request is admitted as soon as the device arena
 * (16q + 8pq + 4mq + 8qS bytes) and the pinned host arena (16q + 8pq + 8qS
 * bytes) have room for it; its sites then share kriging launches with the
 * sites of every other admitted request (at most batch_sites per launch,
 * shortest remaining request first, requests older than 0.25 s first), and
 * several launches and result copies are in flight at once.
 *
 * The standard normals of a request are z(i, s) = Philox4x32-10 of counter
 * (i, s, 0, 0) under key `seed`, turned into a normal by Box-Muller (see
 * philox_normal in the library sources), so a request's result is the same
 * whether it runs alone or among other requests.  Its neighbors are the m
 * nearest fitted sites (nngp_nn2).  The result has the layout of p.y.0:
 * q x nSamples column-major.
 *
 * Timestamps are CLOCK_MONOTONIC seconds.
 * --------------------------------------------------------------------- */

typedef struct nngp_server nngp_server;

typedef struct nngp_server_opts {
  unsigned long long device_bytes; /* request arena in HBM (0: 8 GiB)         */
  unsigned long long host_bytes;   /* pinned host arena (0: 4 GiB)            */
  int batch_sites;                 /* most new sites per launch (0: 16384)    */
  int depth;                       /* launches in flight, 1..8 (0: 3)         */
  int device;                      /* HIP device ordinal                      */
  int reserved[8];
} nngp_server_opts;

typedef struct nngp_request_info {
  int q, nSamples;
  int status;                 /* NNGP_OK or the request's error           */
  double t_submit, t_admit, t_done;
  const double *y0;           /* q x nSamples result in the server's pinned
                                 arena, valid until nngp_server_release;
                                 NULL when status != NNGP_OK               */
} nngp_request_info;

typedef struct nngp_server_stats {
  long long submitted, completed, launches;
  int pending, active;        /* waiting for room / admitted, unfinished  */
  double site_samples;        /* q * nSamples summed over finished requests */
  unsigned long long device_bytes_used, device_bytes_peak;
  unsigned long long host_bytes_used, host_bytes_peak;
  double kernel_ms;           /* device time of the retired kriging launches */
  double kernel_site_samples; /* their sites x nSamples                      */
  long long done_unreported;  /* finished requests no wait has handed out   */
} nngp_server_stats;

/* X n x p and coords n x 2 column-major (the fitted rows in any fixed
 * order), y n, betaSamples p x nSamples, thetaSamples nTheta x nSamples
 * (the p.beta.samples / p.theta.samples columns of a fit). */
int nngp_server_create(nngp_server **out, const double *X, const double *y,
                       const double *coords, int n, int p, int m,
                       const double *betaSamples, const double *thetaSamples,
                       int nSamples, int covModel,
                       const nngp_server_opts *opts);
/* Cancels requests not yet admitted (status NNGP_ERR_INTERRUPT), lets
 * admitted ones finish, then frees everything. */
void nngp_server_destroy(nngp_server *srv);
/* Queue a request of q new sites (X0 q x p, coords0 q x 2, column-major;
 * both copied before returning).  Fails with NNGP_ERR_NOMEM when the request
 * alone exceeds an arena. */
int nngp_server_submit(nngp_server *srv, int q, const double *X0,
                       const double *coords0, unsigned long long seed,
                       long long *ticket);
/* Block until request *ticket finishes, or, with *ticket == -1, until any
 * finished request not yet reported (its ticket is stored in *ticket).
 * timeout_s < 0 waits without limit; on expiry NNGP_ERR_TIMEOUT. */
int nngp_server_wait(nngp_server *srv, long long *ticket, double timeout_s,
                     nngp_request_info *info);
/* Give the finished request's result memory back to the server. */
int nngp_server_release(nngp_server *srv, long long ticket);
/* submit + wait + copy the result into pY0 (q x nSamples) + release. */
int nngp_server_predict(nngp_server *srv, int q, const double *X0,
                        const double *coords0, unsigned long long seed,
                        double *pY0);
int nngp_server_get_stats(nngp_server *srv, nngp_server_stats *st);

/* ------------------------------------------------------------------------
 * Conjugate-model k-fold grid search: the loop of R/spConjNNGP.R:287-357
 * (one cNNGP call per fold, src/cNNGP.cpp:73-358, then the CRPS / RMSPE
 * scores of R/util.R:247-255) as one batched, graph-captured run over all
 * k folds x g grid points.
 *
 * nngp_cvgrid_create fixes the folds: hoIndx is nHo x kFold column-major
 * (fold f's held-out rows at hoIndx[f*nHo ...], 0-based, the R code's
 * `ho.id - 1`); each fold trains on the other n - nHo rows in their given
 * (NNGP) order with its own neighbor index (the mkNNIndxCB of
 * R/spConjNNGP.R:296-300) and predicts its held-out rows from their
 * n.neighbors nearest training rows (the nn2 of R/spConjNNGP.R:311).
 *
 * nngp_cvgrid_run scores g grid points (thetaAlpha nTheta x g, the
 * {phi, alpha[, nu]} columns cNNGP takes).  scores is g x 2 column-major,
 * columns rmspe and crps: the k.fold.scores of R/spConjNNGP.R:350 (each
 * the mean over folds of the per-fold score).  The launch shape of the
 * updateConjBF pass is chosen by timing the candidates on the first run
 * (sites per wavefront 64 or 1, block / LDS tile size); a different choice
 * changes only the order of floating-point sums.
 * --------------------------------------------------------------------- */

typedef struct nngp_cvgrid nngp_cvgrid;

typedef struct nngp_cvgrid_detail {
  /* optional outputs (NULL: not returned), problem j = fold * g + point: */
  double *betaHat;            /* p x (k g)                                  */
  double *ab;                 /* 2 x (k g): a, b                            */
  double *bBInv;              /* p*p x (k g): (X'V^-1X)^-1 of the fold      */
  double *y0Hat, *y0HatVar;   /* nHo x (k g)                                */
  /* filled by every run: */
  double graph_ms;            /* device time of the run's graph             */
  int bf_family, bf_threads, bf_mode;  /* launch shape in use, as in
                                          nngp_tune_report                */
  int tune_n;                 /* candidates timed at the last tuning        */
  int tune_family[NNGP_TUNE_MAX], tune_threads[NNGP_TUNE_MAX],
      tune_mode[NNGP_TUNE_MAX];
  double tune_ms[NNGP_TUNE_MAX];
} nngp_cvgrid_detail;

int nngp_cvgrid_create(nngp_cvgrid **out, const double *y, const double *X,
                       int p, int n, const double *coords, int m, int kFold,
                       const int *hoIndx, int nHo, const double *sigmaSqIG,
                       int covModel, const nngp_opts *opts);
int nngp_cvgrid_run(nngp_cvgrid *cv, const double *thetaAlpha, int g,
                    double *scores, nngp_cvgrid_detail *detail);
void nngp_cvgrid_destroy(nngp_cvgrid *cv);

#ifdef __cplusplus
}
#endif

#endif /* NNGP_H */
