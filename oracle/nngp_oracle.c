/*
 * nngp_oracle.c — CPU oracle for spNNGP 1.0.1's NNGP hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see nngp_oracle.h).  Written from the
 * reference's algorithms; compiled with -ffp-contract=off so that every
 * product/sum rounds separately, as the reference's x86-64 build does.
 */
#include "nngp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static _Thread_local char g_err[256];

const char *or_last_error(void) { return g_err; }

static int fail(int code, const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

/* ------------------------------------------------------------------ */
/* layout and geometry                                                 */
/* ------------------------------------------------------------------ */

/* src/util.cpp:55-70: offset and count of site i in the flat arrays.  The
 * reference evaluates the offsets in double and truncates; the products
 * are exact integers in double for every n that fits in int32. */
void or_getNNIndx(int i, int m, int *iNNIndx, int *iNN) {
  if (i == 0) {
    *iNNIndx = 0;
    *iNN = 0;
  } else if (i < m) {
    *iNNIndx = (int)((double)i / 2 * (i - 1));
    *iNN = i;
  } else {
    *iNNIndx = (int)((double)m / 2 * (m - 1) + (double)(i - m) * m);
    *iNN = m;
  }
}

static long long n_indx(int n, int m) {
  /* src/rNNGP.cpp:218 */
  return (long long)((double)(1 + m) / 2 * m + (double)(n - m - 1) * m);
}

/* src/util.cpp:51-53: sqrt(pow(a1-b1,2)+pow(a2-b2,2)); pow(x,2) == x*x. */
double or_dist2(double a1, double a2, double b1, double b2) {
  double dx = a1 - b1, dy = a2 - b2;
  return sqrt(dx * dx + dy * dy);
}

/* ------------------------------------------------------------------ */
/* Bessel K_nu (Temme 1975: series for x <= 2, Steed continued fraction  */
/* CF2 for x > 2, then upward recurrence in the order)                 */
/* ------------------------------------------------------------------ */

/* Coefficients of 1/Gamma(z) = sum_k c_k z^k (Abramowitz & Stegun 6.1.34),
 * even k: used for Gamma_1(mu) = -(c2 + c4 mu^2 + c6 mu^4 + ...) where the
 * difference formula cancels. */
static const double RG_C2 = 0.5772156649015329;
static const double RG_C4 = -0.0420026350340952;
static const double RG_C6 = -0.0421977345555443;
static const double RG_C8 = 0.0072189432466630;
static const double RG_C10 = -0.0002152416741149;
static const double RG_C12 = -0.0000201348547807;
static const double RG_C14 = 0.0000011330272320;

static void temme_gammas(double mu, double *g1, double *g2, double *rgp,
                         double *rgm) {
  double rp = 1.0 / tgamma(1.0 + mu); /* 1/Gamma(1+mu) */
  double rm = 1.0 / tgamma(1.0 - mu); /* 1/Gamma(1-mu) */
  *rgp = rp;
  *rgm = rm;
  *g2 = 0.5 * (rm + rp);
  if (fabs(mu) < 0.1) {
    double u = mu * mu;
    *g1 = -(RG_C2 +
            u * (RG_C4 +
                 u * (RG_C6 + u * (RG_C8 + u * (RG_C10 + u * (RG_C12 +
                                                            u * RG_C14))))));
  } else {
    *g1 = (rm - rp) / (2.0 * mu);
  }
}

double or_bessel_k(double x, double nu) {
  const double eps = 1e-17;
  const double pi = 3.14159265358979323846;
  if (!(x > 0.0)) return x == 0.0 ? INFINITY : NAN;
  nu = fabs(nu);
  int nl = (int)(nu + 0.5);
  double mu = nu - nl; /* in [-1/2, 1/2) */
  double kmu, kmu1;    /* K_mu, K_{mu+1} */
  if (x <= 2.0) {
    double g1, g2, rgp, rgm;
    temme_gammas(mu, &g1, &g2, &rgp, &rgm);
    double half = 0.5 * x;
    double lg = -log(half); /* ln(2/x) */
    double sg = mu * lg;    /* sigma */
    double pm = pi * mu;
    double fa = fabs(pm) < 1e-300 ? 1.0 : pm / sin(pm);
    double fs = fabs(sg) < 1e-300 ? 1.0 : sinh(sg) / sg;
    double f = fa * (g1 * cosh(sg) + g2 * fs * lg);
    double es = exp(sg);
    double pk = 0.5 * es / rgp; /* p_0 = (x/2)^-mu Gamma(1+mu) / 2 */
    double qk = 0.5 / (es * rgm); /* q_0 = (x/2)^mu Gamma(1-mu) / 2 */
    double ck = 1.0;
    double d = half * half;
    double s0 = f, s1 = pk;
    for (int k = 1; k < 500; k++) {
      f = (k * f + pk + qk) / ((double)k * k - mu * mu);
      ck *= d / k;
      pk /= (k - mu);
      qk /= (k + mu);
      double t0 = ck * f;
      s0 += t0;
      s1 += ck * (pk - k * f);
      if (fabs(t0) < fabs(s0) * eps) break;
    }
    kmu = s0;
    kmu1 = s1 * (2.0 / x);
  } else {
    /* Steed's algorithm for CF2 */
    double b = 2.0 * (1.0 + x);
    double d = 1.0 / b;
    double h = d, dh = d;
    double q1 = 0.0, q2 = 1.0;
    double a1 = 0.25 - mu * mu;
    double q = a1, c = a1;
    double a = -a1;
    double s = 1.0 + q * dh;
    for (int i = 2; i < 100000; i++) {
      a -= 2.0 * (i - 1);
      c = -a * c / i;
      double qn = (q1 - b * q2) / a;
      q1 = q2;
      q2 = qn;
      q += c * qn;
      b += 2.0;
      d = 1.0 / (b + a * d);
      dh = (b * d - 1.0) * dh;
      h += dh;
      double ds = q * dh;
      s += ds;
      if (fabs(ds / s) < eps) break;
    }
    h = a1 * h;
    kmu = sqrt(pi / (2.0 * x)) * exp(-x) / s;
    kmu1 = kmu * (mu + x + 0.5 - h) / x;
  }
  for (int i = 1; i <= nl; i++) {
    double kn = (mu + i) * (2.0 / x) * kmu1 + kmu;
    kmu = kmu1;
    kmu1 = kn;
  }
  return kmu;
}

/* src/util.cpp:208-245 */
double or_spCor(double D, double phi, double nu, int covModel) {
  if (covModel == 0) {
    return exp(-phi * D);
  } else if (covModel == 1) {
    if (D > 0 && D <= 1.0 / phi) {
      return 1.0 - 1.5 * phi * D + 0.5 * pow(phi * D, 3);
    } else if (D >= 1.0 / phi) {
      return 0.0;
    } else {
      return 1.0;
    }
  } else if (covModel == 2) {
    if (D * phi > 0.0) {
      return pow(D * phi, nu) / (pow(2, nu - 1) * tgamma(nu)) *
             or_bessel_k(D * phi, nu);
    }
    return 1.0;
  } else if (covModel == 3) {
    double t = phi * D;
    return exp(-1.0 * (t * t));
  }
  return NAN;
}

/* ------------------------------------------------------------------ */
/* sorting and neighbor search                                         */
/* ------------------------------------------------------------------ */

/* comparison with NaN last (R's rcmp with nalast = TRUE) */
static int rcmp_last(double x, double y) {
  int nx = isnan(x), ny = isnan(y);
  if (nx && ny) return 0;
  if (nx) return 1;
  if (ny) return -1;
  if (x < y) return -1;
  if (x > y) return 1;
  return 0;
}

/* R's rsort_with_index: shell sort over the gap sequence 1, 4, 13, 40, ...
 * starting from the first gap above n/9. */
void or_rsort_with_index(double *x, int *indx, int n) {
  int h = 1;
  while (h <= n / 9) h = 3 * h + 1;
  for (; h > 0; h /= 3) {
    for (int i = h; i < n; i++) {
      double v = x[i];
      int iv = indx[i];
      int j = i;
      while (j >= h && rcmp_last(x[j - h], v) > 0) {
        x[j] = x[j - h];
        indx[j] = indx[j - h];
        j -= h;
      }
      x[j] = v;
      indx[j] = iv;
    }
  }
}

static int check_nn_args(int n, int m) {
  if (n < 1 || m < 1) return fail(OR_ERR_ARG, "n and m must be positive");
  if (m >= n) return fail(OR_ERR_ARG, "m must be smaller than n");
  return OR_OK;
}

/* src/nn.cpp:63-111 */
int or_mkNNIndx(int n, int m, const double *coords, int *nnIndx,
                double *nnDist, int *nnIndxLU, int nThreads) {
  int st = check_nn_args(n, m);
  if (st) return st;
  long long nIndx = n_indx(n, m);
  for (long long i = 0; i < nIndx; i++) {
    nnDist[i] = INFINITY;
    nnIndx[i] = 0;
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nThreads > 0 ? nThreads : 1)
#endif
  for (int i = 0; i < n; i++) {
    int off, cnt;
    or_getNNIndx(i, m, &off, &cnt);
    nnIndxLU[i] = off;
    nnIndxLU[n + i] = cnt;
    for (int j = 0; j < i; j++) {
      double d = or_dist2(coords[i], coords[n + i], coords[j], coords[n + j]);
      if (d < nnDist[off + cnt - 1]) {
        nnDist[off + cnt - 1] = d;
        nnIndx[off + cnt - 1] = j;
        or_rsort_with_index(&nnDist[off], &nnIndx[off], cnt);
      }
    }
  }
  return OR_OK;
}

/* src/nn.cpp:128-134 */
static double cb_dmi(const double *x, const double *c, int inc) {
  double t = x[0] + x[inc] - c[0] - c[inc];
  return t * t;
}
static double cb_dei(const double *x, const double *c, int inc) {
  double a = x[0] - c[0], b = x[inc] - c[inc];
  return a * a + b * b;
}

/* src/nn.cpp:136-211.  A site with no predecessors (m == 0) has nothing to
 * fill; the reference reads rSNNDist[-1] there and writes nothing. */
static void cb_fastNN(int m, int n, const double *coords, int ui,
                      const int *sIndx, int *rSIndx, double *rSNNDist) {
  if (m == 0) return;
  for (int k = 0; k < m; k++) rSNNDist[k] = INFINITY;
  int i = ui, j = ui;
  int up = 1, down = 1;
  const double *q = &coords[sIndx[ui]];
  while (up || down) {
    if (i == 0) down = 0;
    if (j == n - 1) up = 0;
    if (down) {
      i--;
      double dm = cb_dmi(q, &coords[sIndx[i]], n);
      if (dm > 2 * rSNNDist[m - 1]) {
        down = 0;
      } else {
        double de = cb_dei(q, &coords[sIndx[i]], n);
        if (de < rSNNDist[m - 1] && sIndx[i] < sIndx[ui]) {
          rSNNDist[m - 1] = de;
          rSIndx[m - 1] = sIndx[i];
          or_rsort_with_index(rSNNDist, rSIndx, m);
        }
      }
    }
    if (up) {
      j++;
      double dm = cb_dmi(q, &coords[sIndx[j]], n);
      if (dm > 2 * rSNNDist[m - 1]) {
        up = 0;
      } else {
        double de = cb_dei(q, &coords[sIndx[j]], n);
        if (de < rSNNDist[m - 1] && sIndx[j] < sIndx[ui]) {
          rSNNDist[m - 1] = de;
          rSIndx[m - 1] = sIndx[j];
          or_rsort_with_index(rSNNDist, rSIndx, m);
        }
      }
    }
  }
  for (int k = 0; k < m; k++) rSNNDist[k] = sqrt(rSNNDist[k]);
}

/* src/nn.cpp:214-257 */
int or_mkNNIndxCB(int n, int m, const double *coords, int *nnIndx,
                  double *nnDist, int *nnIndxLU, int nThreads) {
  int st = check_nn_args(n, m);
  if (st) return st;
  long long nIndx = n_indx(n, m);
  for (long long i = 0; i < nIndx; i++) {
    nnDist[i] = 0.0;
    nnIndx[i] = 0;
  }
  int *sIndx = (int *)malloc(sizeof(int) * (size_t)n);
  double *u = (double *)malloc(sizeof(double) * (size_t)n);
  if (!sIndx || !u) {
    free(sIndx);
    free(u);
    return fail(OR_ERR_ARG, "out of memory");
  }
  for (int i = 0; i < n; i++) {
    sIndx[i] = i;
    u[i] = coords[i] + coords[n + i];
  }
  or_rsort_with_index(u, sIndx, n);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nThreads > 0 ? nThreads : 1)
#endif
  for (int i = 0; i < n; i++) {
    int off, cnt;
    or_getNNIndx(sIndx[i], m, &off, &cnt);
    nnIndxLU[sIndx[i]] = off;
    nnIndxLU[n + sIndx[i]] = cnt;
    cb_fastNN(cnt, n, coords, i, sIndx, &nnIndx[off], &nnDist[off]);
  }
  free(sIndx);
  free(u);
  return OR_OK;
}

/* RANN nn2 (eps = 0): plain scan, ties by ascending index. */
int or_nn2(int n, const double *coords, int q, const double *coords0, int k,
           int *nnIndx0, double *nnDist0) {
  if (k < 1 || k > n) return fail(OR_ERR_ARG, "k must be in 1..n");
  double *bd = (double *)malloc(sizeof(double) * (size_t)k);
  int *bi = (int *)malloc(sizeof(int) * (size_t)k);
  for (int i = 0; i < q; i++) {
    for (int t = 0; t < k; t++) {
      bd[t] = INFINITY;
      bi[t] = -1;
    }
    for (int j = 0; j < n; j++) {
      double dx = coords0[i] - coords[j], dy = coords0[q + i] - coords[n + j];
      double d2 = dx * dx + dy * dy;
      if (d2 < bd[k - 1]) {
        int t = k - 1;
        while (t > 0 && bd[t - 1] > d2) {
          bd[t] = bd[t - 1];
          bi[t] = bi[t - 1];
          t--;
        }
        bd[t] = d2;
        bi[t] = j;
      }
    }
    for (int t = 0; t < k; t++) {
      nnIndx0[i + (size_t)q * t] = bi[t];
      if (nnDist0) nnDist0[i + (size_t)q * t] = sqrt(bd[t]);
    }
  }
  free(bd);
  free(bi);
  return OR_OK;
}

/* ------------------------------------------------------------------ */
/* dense kernels: LAPACK dpotf2 / dtrti2 / dlauu2, BLAS dsymv, dtrmv   */
/* (reference implementations' loop orders, lower triangle, lda = n)   */
/* ------------------------------------------------------------------ */

#define A_(i, j) A[(size_t)(j) * n + (i)]

/* dpotf2, UPLO = 'L': returns 0 or the 1-based failing column. */
int or_potrf_L(int n, double *A) {
  for (int j = 0; j < n; j++) {
    double ajj = A_(j, j);
    double dot = 0.0;
    for (int k = 0; k < j; k++) dot += A_(j, k) * A_(j, k);
    ajj = ajj - dot;
    if (ajj <= 0.0 || isnan(ajj)) {
      A_(j, j) = ajj;
      return j + 1;
    }
    ajj = sqrt(ajj);
    A_(j, j) = ajj;
    if (j < n - 1) {
      /* dgemv('N'): A(j+1:n, j) -= A(j+1:n, 0:j) * A(j, 0:j)^T */
      for (int k = 0; k < j; k++) {
        double t = -A_(j, k);
        for (int i = j + 1; i < n; i++) A_(i, j) += t * A_(i, k);
      }
      double r = 1.0 / ajj;
      for (int i = j + 1; i < n; i++) A_(i, j) *= r;
    }
  }
  return 0;
}

/* dtrti2('L','N') followed by dlauu2('L'): inv(A) from its lower Cholesky
 * factor, lower triangle. */
int or_potri_L(int n, double *A) {
  for (int j = 0; j < n; j++)
    if (A_(j, j) == 0.0) return j + 1;
  /* dtrti2 lower, non-unit */
  for (int j = n - 1; j >= 0; j--) {
    A_(j, j) = 1.0 / A_(j, j);
    double ajj = -A_(j, j);
    if (j < n - 1) {
      /* dtrmv('L','N','N') on x = A(j+1:n, j) with T = A(j+1:n, j+1:n) */
      int nn = n - j - 1;
      for (int jj = nn - 1; jj >= 0; jj--) {
        double t = A_(j + 1 + jj, j);
        if (t != 0.0) {
          for (int ii = nn - 1; ii > jj; ii--)
            A_(j + 1 + ii, j) += t * A_(j + 1 + ii, j + 1 + jj);
          A_(j + 1 + jj, j) = t * A_(j + 1 + jj, j + 1 + jj);
        }
      }
      for (int ii = 0; ii < nn; ii++) A_(j + 1 + ii, j) *= ajj;
    }
  }
  /* dlauu2 lower: A := L^T L */
  for (int i = 0; i < n; i++) {
    double aii = A_(i, i);
    if (i < n - 1) {
      double dot = 0.0;
      for (int k = i; k < n; k++) dot += A_(k, i) * A_(k, i);
      A_(i, i) = dot;
      /* dgemv('T'): A(i, 0:i) = aii*A(i,0:i) + A(i+1:n, 0:i)^T A(i+1:n, i) */
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

/* dsymv('L'), alpha = 1, beta = 0 */
void or_symv_L(int n, const double *A, const double *x, double *y) {
  for (int i = 0; i < n; i++) y[i] = 0.0;
  for (int j = 0; j < n; j++) {
    double t1 = x[j], t2 = 0.0;
    y[j] += t1 * A_(j, j);
    for (int i = j + 1; i < n; i++) {
      y[i] += t1 * A_(i, j);
      t2 += A_(i, j) * x[i];
    }
    y[j] += t2;
  }
}

/* dtrmv('L','N','N') */
static void trmv_L(int n, const double *A, double *x) {
  for (int j = n - 1; j >= 0; j--) {
    double t = x[j];
    if (t != 0.0) {
      for (int i = n - 1; i > j; i--) x[i] += t * A_(i, j);
      x[j] = t * A_(j, j);
    }
  }
}

#undef A_

static double ddot(int n, const double *x, int incx, const double *y,
                   int incy) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += x[(size_t)i * incx] * y[(size_t)i * incy];
  return s;
}

/* ------------------------------------------------------------------ */
/* updateBF / updateConjBF / Q                                         */
/* ------------------------------------------------------------------ */

/* One site's neighbor system: C = s2 R(N,N) + t2 I (lower), c = s2 R(N,i),
 * potrf -> potri -> symv, F = s2 - B.c + t2  (src/rNNGP.cpp:74-89). */
static int site_bf(int i, int n, int mi, const int *nb, const double *coords,
                   double s2, double t2, double phi, double nu, int covModel,
                   double *C, double *c, double *Bout, double *Fout) {
  for (int k = 0; k < mi; k++) {
    int a = nb[k];
    double e = or_dist2(coords[i], coords[n + i], coords[a], coords[n + a]);
    c[k] = s2 * or_spCor(e, phi, nu, covModel);
    for (int l = 0; l <= k; l++) {
      int b = nb[l];
      e = or_dist2(coords[a], coords[n + a], coords[b], coords[n + b]);
      C[l * mi + k] = s2 * or_spCor(e, phi, nu, covModel);
      if (l == k) C[l * mi + k] += t2;
    }
  }
  if (or_potrf_L(mi, C) != 0) return OR_ERR_NOT_PD;
  if (or_potri_L(mi, C) != 0) return OR_ERR_NOT_PD;
  or_symv_L(mi, C, c, Bout);
  *Fout = s2 - ddot(mi, Bout, 1, c, 1) + t2;
  return OR_OK;
}

static int bf_common(double *B, double *F, const double *coords,
                     const int *nnIndx, const int *nnIndxLU, int n, int m,
                     double s2, double t2, double phi, double nu,
                     int covModel, int nThreads) {
  if (covModel < 0 || covModel > 3)
    return fail(OR_ERR_COV_MODEL,
                "c++ Rf_error: cov.model is not correctly specified");
  int bad = -1;
  int nt = nThreads > 0 ? nThreads : 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
  {
    double *C = (double *)malloc(sizeof(double) * (size_t)m * m);
    double *c = (double *)malloc(sizeof(double) * (size_t)m);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int i = 0; i < n; i++) {
      if (i > 0) {
        int mi = nnIndxLU[n + i];
        if (site_bf(i, n, mi, &nnIndx[nnIndxLU[i]], coords, s2, t2, phi, nu,
                    covModel, C, c, &B[nnIndxLU[i]], &F[i]) != OR_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
          if (bad < 0 || i < bad) bad = i;
        }
      } else {
        F[i] = s2 + t2;
      }
    }
    free(C);
    free(c);
  }
  if (bad >= 0) {
    snprintf(g_err, sizeof g_err, "c++ Rf_error: dpotrf failed (site %d)",
             bad);
    return OR_ERR_NOT_PD;
  }
  return OR_OK;
}

/* src/rNNGP.cpp:46-101 */
int or_updateBF(double *B, double *F, const double *coords,
                const int *nnIndx, const int *nnIndxLU, int n, int m,
                const double *theta, int covModel, int nThreads,
                double *logDet) {
  double nu = covModel == 2 ? theta[3] : 0.0;
  int st = bf_common(B, F, coords, nnIndx, nnIndxLU, n, m, theta[0],
                     theta[1], theta[2], nu, covModel, nThreads);
  if (st) return st;
  double ld = 0.0;
  for (int i = 0; i < n; i++) ld += log(F[i]);
  *logDet = ld;
  return OR_OK;
}

/* src/cNNGP.cpp:24-69: correlation scale, F = 1 + alpha - B.c */
int or_updateConjBF(double *B, double *F, const double *coords,
                    const int *nnIndx, const int *nnIndxLU, int n, int m,
                    double phi, double alpha, double nu, int covModel,
                    int nThreads) {
  return bf_common(B, F, coords, nnIndx, nnIndxLU, n, m, 1.0, alpha, phi,
                   nu, covModel, nThreads);
}

/* src/util.cpp:261-280 */
double or_Q(const double *B, const double *F, const double *u,
            const double *v, int n, const int *nnIndx, const int *nnIndxLU) {
  double q = 0.0;
  for (int i = 0; i < n; i++) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j < nnIndxLU[n + i]; j++) {
      a += B[nnIndxLU[i] + j] * u[nnIndx[nnIndxLU[i] + j]];
      b += B[nnIndxLU[i] + j] * v[nnIndx[nnIndxLU[i] + j]];
    }
    q += (u[i] - a) * (v[i] - b) / F[i];
  }
  return q;
}

/* ------------------------------------------------------------------ */
/* RNG helpers with R's argument semantics                              */
/* ------------------------------------------------------------------ */

/* R's rnorm(mu, sigma): sigma == 0 or non-finite mu returns mu and draws
 * nothing. */
static double r_rnorm(const or_rng *r, double mu, double sd) {
  if (isnan(mu) || !isfinite(sd) || sd < 0.0) return NAN;
  if (sd == 0.0 || !isfinite(mu)) return mu;
  return mu + sd * r->norm_rand(r->state);
}

/* R's runif(a, b) */
static double r_runif(const or_rng *r, double a, double b) {
  if (!isfinite(a) || !isfinite(b) || b < a) return NAN;
  if (a == b) return a;
  double u;
  do {
    u = r->unif_rand(r->state);
  } while (u <= 0.0 || u >= 1.0);
  return a + (b - a) * u;
}

/* src/util.cpp:43-49 */
static double logit_ab(double t, double a, double b) {
  return log((t - a) / (b - t));
}
static double logit_inv(double z, double a, double b) {
  return b - (b - a) / (1 + exp(z));
}

/* src/util.cpp:28-41: des = mu + L z */
static void mvrnorm(const or_rng *r, double *des, const double *mu,
                    const double *L, int dim) {
  for (int i = 0; i < dim; i++) des[i] = r_rnorm(r, 0, 1);
  trmv_L(dim, L, des);
  for (int i = 0; i < dim; i++) des[i] += mu[i];
}

void or_nngp_draw(const double *B, const double *F, const int *nnIndx,
                  const int *nnIndxLU, int n, const double *z, double *w) {
  for (int i = 0; i < n; i++) {
    if (i == 0) {
      w[i] = sqrt(F[i]) * z[i];
    } else {
      double s = 0.0;
      for (int j = 0; j < nnIndxLU[n + i]; j++)
        s += B[nnIndxLU[i] + j] * w[nnIndx[nnIndxLU[i] + j]];
      w[i] = s + sqrt(F[i]) * z[i];
    }
  }
}

/* src/rNNGPReplicated.cpp:78-230 (the sample loop at :181-199) */
int or_rNNGPReplicated(const double *X, int p, int n, int m,
                       const double *coords, int covModel, const int *nnIndx,
                       const int *nnIndxLU, const double *beta,
                       const double *theta, int nSamples, int nThreads,
                       const or_rng *rng, double *yRep) {
  const int nTheta = covModel == 2 ? 4 : 3;
  const long long nIndx = (long long)(1 + m) * m / 2 + (long long)(n - m - 1) * m;
  double *B = (double *)malloc(sizeof(double) * (size_t)(nIndx > 0 ? nIndx : 1));
  double *F = (double *)malloc(sizeof(double) * (size_t)n);
  double *z = (double *)malloc(sizeof(double) * (size_t)n);
  if (!B || !F || !z) {
    free(B);
    free(F);
    free(z);
    return fail(OR_ERR_ARG, "allocation failed");
  }
  int st = OR_OK;
  for (int s = 0; s < nSamples && st == OR_OK; s++) {
    double ld;
    st = or_updateBF(B, F, coords, nnIndx, nnIndxLU, n, m,
                     theta + (size_t)s * nTheta, covModel, nThreads, &ld);
    if (st) break;
    for (int i = 0; i < n; i++) z[i] = r_rnorm(rng, 0.0, 1.0);
    double *v = yRep + (size_t)s * n;
    or_nngp_draw(B, F, nnIndx, nnIndxLU, n, z, v);
    /* dgemv('N', n, p, 1, X, n, beta + s, nSamples, 1, v, 1) */
    for (int j = 0; j < p; j++) {
      const double bj = beta[s + (size_t)j * nSamples];
      for (int i = 0; i < n; i++) v[i] += X[i + (size_t)j * n] * bj;
    }
  }
  free(B);
  free(F);
  free(z);
  return st;
}

/* ------------------------------------------------------------------ */
/* rNNGP: response-model MCMC                                          */
/* ------------------------------------------------------------------ */

static double log_post(const double *th, double logDet, double Q,
                       const double *sIG, const double *tIG,
                       const double *phiU, const double *nuU, int matern) {
  double lp = -0.5 * logDet - 0.5 * Q;
  lp += log(th[2] - phiU[0]) + log(phiU[1] - th[2]);
  lp += -1.0 * (1.0 + sIG[0]) * log(th[0]) - sIG[1] / th[0] + log(th[0]);
  lp += -1.0 * (1.0 + tIG[0]) * log(th[1]) - tIG[1] / th[1] + log(th[1]);
  if (matern) lp += log(th[3] - nuU[0]) + log(nuU[1] - th[3]);
  return lp;
}

/* src/rNNGP.cpp:105-420 */
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
             double *yRepSamples, int *nAccept) {
  if (covModel < 0 || covModel > 3)
    return fail(OR_ERR_COV_MODEL,
                "c++ Rf_error: cov.model is not correctly specified");
  int matern = covModel == 2;
  int nTheta = matern ? 4 : 3;
  double theta[4] = {sigmaSqStarting, tauSqStarting, phiStarting,
                     matern ? nuStarting : 0.0};
  double tuning[4] = {sigmaSqTuning, tauSqTuning, phiTuning,
                      matern ? nuTuning : 0.0};
  double thetaCand[4] = {0, 0, 0, 0};
  double nuU[2] = {matern ? nuUnif[0] : 0.0, matern ? nuUnif[1] : 0.0};
  long long nIndx = n_indx(n, m);
  double *beta = (double *)calloc((size_t)p, sizeof(double));
  double *B = (double *)calloc((size_t)nIndx, sizeof(double));
  double *F = (double *)calloc((size_t)n, sizeof(double));
  double *tmp_pp = (double *)calloc((size_t)p * p, sizeof(double));
  double *tmp_p = (double *)calloc((size_t)p, sizeof(double));
  double *tmp_p2 = (double *)calloc((size_t)p, sizeof(double));
  double *tmp_n = (double *)calloc((size_t)n, sizeof(double));
  double *tmp_n2 = (double *)calloc((size_t)n, sizeof(double));
  double *z = (double *)calloc((size_t)n, sizeof(double));
  int st = OR_OK, accept = 0, repCnt = 0;
  for (int i = 0; i < p; i++) beta[i] = betaStarting[i];

  double logDetCurrent, logDetCand, QCurrent, QCand;
  st = or_updateBF(B, F, coords, nnIndx, nnIndxLU, n, m, theta, covModel,
                   nThreads, &logDetCurrent);
  if (st) goto done;
  int thetaUpdate = 1;
  for (int s = 0; s < nSamples; s++) {
    if (thetaUpdate) {
      thetaUpdate = 0;
      for (int i = 0; i < p; i++) {
        tmp_p[i] = or_Q(B, F, &X[(size_t)n * i], y, n, nnIndx, nnIndxLU);
        for (int j = 0; j <= i; j++)
          tmp_pp[j * p + i] = or_Q(B, F, &X[(size_t)n * j],
                                   &X[(size_t)n * i], n, nnIndx, nnIndxLU);
      }
      if (or_potrf_L(p, tmp_pp) || or_potri_L(p, tmp_pp)) {
        st = fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed");
        goto done;
      }
      or_symv_L(p, tmp_pp, tmp_p, tmp_p2);
      if (or_potrf_L(p, tmp_pp)) {
        st = fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed");
        goto done;
      }
    }
    mvrnorm(rng, beta, tmp_p2, tmp_pp, p);
    /* dgemv: tmp_n = X beta; then tmp_n -= y */
    for (int i = 0; i < n; i++) tmp_n[i] = 0.0;
    for (int j = 0; j < p; j++) {
      double t = beta[j];
      for (int i = 0; i < n; i++) tmp_n[i] += t * X[(size_t)n * j + i];
    }
    if (nRep && repIndx[s])
      for (int i = 0; i < n; i++)
        yRepSamples[(size_t)repCnt * n + i] = tmp_n[i];
    for (int i = 0; i < n; i++) tmp_n[i] += -1.0 * y[i];

    st = or_updateBF(B, F, coords, nnIndx, nnIndxLU, n, m, theta, covModel,
                     nThreads, &logDetCurrent);
    if (st) goto done;
    if (nRep && repIndx[s]) {
      for (int i = 0; i < n; i++) z[i] = r_rnorm(rng, 0.0, 1.0);
      or_nngp_draw(B, F, nnIndx, nnIndxLU, n, z, tmp_n2);
      for (int i = 0; i < n; i++)
        yRepSamples[(size_t)repCnt * n + i] += tmp_n2[i];
      repCnt++;
    }
    QCurrent = or_Q(B, F, tmp_n, tmp_n, n, nnIndx, nnIndxLU);
    double lpCur = log_post(theta, logDetCurrent, QCurrent, sigmaSqIG,
                            tauSqIG, phiUnif, nuU, matern);

    thetaCand[2] = logit_inv(
        r_rnorm(rng, logit_ab(theta[2], phiUnif[0], phiUnif[1]), tuning[2]),
        phiUnif[0], phiUnif[1]);
    thetaCand[0] = exp(r_rnorm(rng, log(theta[0]), tuning[0]));
    thetaCand[1] = exp(r_rnorm(rng, log(theta[1]), tuning[1]));
    if (matern)
      thetaCand[3] = logit_inv(
          r_rnorm(rng, logit_ab(theta[3], nuU[0], nuU[1]), tuning[3]),
          nuU[0], nuU[1]);

    st = or_updateBF(B, F, coords, nnIndx, nnIndxLU, n, m, thetaCand,
                     covModel, nThreads, &logDetCand);
    if (st) goto done;
    QCand = or_Q(B, F, tmp_n, tmp_n, n, nnIndx, nnIndxLU);
    double lpCand = log_post(thetaCand, logDetCand, QCand, sigmaSqIG,
                             tauSqIG, phiUnif, nuU, matern);
    if (r_runif(rng, 0.0, 1.0) <= exp(lpCand - lpCur)) {
      thetaUpdate = 1;
      for (int k = 0; k < nTheta; k++) theta[k] = thetaCand[k];
      accept++;
    }
    for (int k = 0; k < p; k++) pBetaSamples[(size_t)s * p + k] = beta[k];
    for (int k = 0; k < nTheta; k++)
      pThetaSamples[(size_t)s * nTheta + k] = theta[k];
  }
done:
  if (nAccept) *nAccept = accept;
  free(beta);
  free(B);
  free(F);
  free(tmp_pp);
  free(tmp_p);
  free(tmp_p2);
  free(tmp_n);
  free(tmp_n2);
  free(z);
  return st;
}

/* ------------------------------------------------------------------ */
/* rNNGPPredict                                                         */
/* ------------------------------------------------------------------ */

/* src/rNNGPPredict.cpp:139-183 with one thread: y0[s*q+i] uses z[i*S+s] */
int or_rNNGPPredict(const double *X, const double *y, const double *coords,
                    int n, int p, int m, const double *X0,
                    const double *coords0, int q, const int *nnIndx0,
                    const double *betaSamples, const double *thetaSamples,
                    int nSamples, int covModel, int nThreads, const double *z,
                    double *pY0) {
  if (covModel < 0 || covModel > 3)
    return fail(OR_ERR_COV_MODEL,
                "c++ Rf_error: cov.model is not correctly specified");
  int nTheta = covModel == 2 ? 4 : 3;
  int bad = 0;
  int nt = nThreads > 0 ? nThreads : 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
  {
    double *C = (double *)malloc(sizeof(double) * (size_t)m * m);
    double *c = (double *)malloc(sizeof(double) * (size_t)m);
    double *w = (double *)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < q; i++) {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int s = 0; s < nSamples; s++) {
        const double *th = &thetaSamples[(size_t)s * nTheta];
        double s2 = th[0], t2 = th[1], phi = th[2];
        double nu = covModel == 2 ? th[3] : 0.0;
        for (int k = 0; k < m; k++) {
          int a = nnIndx0[i + (size_t)q * k];
          double d = or_dist2(coords[a], coords[n + a], coords0[i],
                              coords0[q + i]);
          c[k] = s2 * or_spCor(d, phi, nu, covModel);
          for (int l = 0; l < m; l++) {
            int b = nnIndx0[i + (size_t)q * l];
            d = or_dist2(coords[a], coords[n + a], coords[b], coords[n + b]);
            C[l * m + k] = s2 * or_spCor(d, phi, nu, covModel);
            if (k == l) C[l * m + k] += t2;
          }
        }
        if (or_potrf_L(m, C) || or_potri_L(m, C)) {
          bad = 1;
          continue;
        }
        or_symv_L(m, C, c, w);
        const double *bs = &betaSamples[(size_t)s * p];
        double d = 0.0;
        for (int k = 0; k < m; k++) {
          int a = nnIndx0[i + (size_t)q * k];
          d += w[k] * (y[a] - ddot(p, &X[a], n, bs, 1));
        }
        pY0[(size_t)s * q + i] =
            ddot(p, &X0[i], q, bs, 1) + d +
            sqrt(s2 + t2 - ddot(m, w, 1, c, 1)) * z[(size_t)i * nSamples + s];
      }
    }
    free(C);
    free(c);
    free(w);
  }
  if (bad) return fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed");
  return OR_OK;
}

/* ------------------------------------------------------------------ */
/* cNNGP                                                                */
/* ------------------------------------------------------------------ */

/* src/cNNGP.cpp:73-358 */
int or_cNNGP(const double *y, const double *X, int p, int n,
             const double *coords, const double *thetaAlpha,
             const int *nnIndx, const int *nnIndxLU, const double *X0,
             const double *coords0, int n0, const int *nnIndx0, int g, int m,
             const double *sigmaSqIG, int covModel, int nThreads,
             double *betaHat, double *ab, double *bBInv, double *bb,
             double *y0Hat, double *y0HatVar) {
  if (covModel < 0 || covModel > 3)
    return fail(OR_ERR_COV_MODEL,
                "c++ Rf_error: cov.model is not correctly specified");
  int nTheta = covModel == 2 ? 3 : 2;
  long long nIndx = n_indx(n, m);
  double *B = (double *)calloc((size_t)nIndx, sizeof(double));
  double *F = (double *)calloc((size_t)n, sizeof(double));
  double *tmp_pp = (double *)calloc((size_t)p * p, sizeof(double));
  double *tmp_p = (double *)calloc((size_t)p, sizeof(double));
  int st = OR_OK;
  int nt = nThreads > 0 ? nThreads : 1;
  for (int k = 0; k < g; k++) {
    double phi = thetaAlpha[k * nTheta];
    double alpha = thetaAlpha[k * nTheta + 1];
    double nu = covModel == 2 ? thetaAlpha[k * nTheta + 2] : 0.0;
    st = or_updateConjBF(B, F, coords, nnIndx, nnIndxLU, n, m, phi, alpha,
                         nu, covModel, nThreads);
    if (st) break;
    for (int i = 0; i < p; i++) {
      tmp_p[i] = or_Q(B, F, &X[(size_t)n * i], y, n, nnIndx, nnIndxLU);
      for (int j = 0; j <= i; j++)
        tmp_pp[j * p + i] = or_Q(B, F, &X[(size_t)n * j], &X[(size_t)n * i],
                                 n, nnIndx, nnIndxLU);
    }
    if (or_potrf_L(p, tmp_pp) || or_potri_L(p, tmp_pp)) {
      st = fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed");
      break;
    }
    double *bk = &betaHat[(size_t)k * p];
    or_symv_L(p, tmp_pp, tmp_p, bk);
    double a = sigmaSqIG[0] + 1.0 * n / 2.0;
    double b = sigmaSqIG[1] +
               0.5 * (or_Q(B, F, y, y, n, nnIndx, nnIndxLU) -
                      ddot(p, bk, 1, tmp_p, 1));
    ab[k * 2] = a;
    ab[k * 2 + 1] = b;
    for (int i = 0; i < p; i++)
      for (int j = 0; j < p; j++)
        bBInv[(size_t)k * p * p + j * p + i] =
            i >= j ? tmp_pp[j * p + i] : tmp_pp[i * p + j];
    for (int i = 0; i < p; i++) bb[(size_t)k * p + i] = tmp_p[i];

    int bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
      double *C0 = (double *)malloc(sizeof(double) * (size_t)m * m);
      double *c0 = (double *)malloc(sizeof(double) * (size_t)m);
      double *w = (double *)malloc(sizeof(double) * (size_t)m);
      double *tm = (double *)malloc(sizeof(double) * (size_t)m);
      double *tmp = (double *)malloc(sizeof(double) * (size_t)m * p);
      double *u = (double *)malloc(sizeof(double) * (size_t)p);
      double *u2 = (double *)malloc(sizeof(double) * (size_t)p);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int i = 0; i < n0; i++) {
        for (int j = 0; j < m; j++) {
          int aj = nnIndx0[(size_t)j * n0 + i];
          double e = or_dist2(coords0[i], coords0[n0 + i], coords[aj],
                              coords[n + aj]);
          c0[j] = or_spCor(e, phi, nu, covModel);
          for (int l = 0; l <= j; l++) {
            int al = nnIndx0[(size_t)l * n0 + i];
            e = or_dist2(coords[aj], coords[n + aj], coords[al],
                         coords[n + al]);
            C0[l * m + j] = or_spCor(e, phi, nu, covModel);
            if (l == j) C0[l * m + j] += alpha;
          }
        }
        if (or_potrf_L(m, C0) || or_potri_L(m, C0)) {
          bad = 1;
          continue;
        }
        or_symv_L(m, C0, c0, w);
        for (int j = 0; j < m; j++) {
          int aj = nnIndx0[(size_t)j * n0 + i];
          tm[j] = y[aj] - ddot(p, &X[aj], n, bk, 1);
        }
        y0Hat[(size_t)k * n0 + i] =
            ddot(p, &X0[i], n0, bk, 1) + ddot(m, w, 1, tm, 1);
        for (int j = 0; j < m; j++) {
          int aj = nnIndx0[(size_t)j * n0 + i];
          for (int t = 0; t < p; t++)
            tmp[(size_t)t * m + j] = X[(size_t)t * n + aj];
        }
        for (int t = 0; t < p; t++) {
          double acc = 0.0;
          for (int j = 0; j < m; j++) acc += tmp[(size_t)t * m + j] * w[j];
          u[t] = acc;
        }
        for (int t = 0; t < p; t++) u[t] = X0[(size_t)t * n0 + i] - u[t];
        or_symv_L(p, tmp_pp, u, u2);
        y0HatVar[(size_t)k * n0 + i] =
            b * (ddot(p, u, 1, u2, 1) + 1.0 + alpha - ddot(m, w, 1, c0, 1)) /
            (a - 1.0);
      }
      free(C0);
      free(c0);
      free(w);
      free(tm);
      free(tmp);
      free(u);
      free(u2);
    }
    if (bad) {
      st = fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed");
      break;
    }
  }
  free(B);
  free(F);
  free(tmp_pp);
  free(tmp_p);
  return st;
}

/* ------------------------------------------------------------------------
 * Philox4x32-10 and the kriging server's normals (test infrastructure for
 * the server; the reference itself draws z from R's RNG)
 * --------------------------------------------------------------------- */
void or_philox4x32_10(unsigned int c[4], unsigned int k0, unsigned int k1) {
  for (int r = 0; r < 10; r++) {
    unsigned long long lo = (unsigned long long)0xD2511F53u * c[0];
    unsigned long long hi = (unsigned long long)0xCD9E8D57u * c[2];
    unsigned int x0 = (unsigned int)(hi >> 32) ^ c[1] ^ k0;
    unsigned int x1 = (unsigned int)hi;
    unsigned int x2 = (unsigned int)(lo >> 32) ^ c[3] ^ k1;
    unsigned int x3 = (unsigned int)lo;
    c[0] = x0;
    c[1] = x1;
    c[2] = x2;
    c[3] = x3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

double or_philox_normal(unsigned long long seed, unsigned int i,
                        unsigned int s) {
  unsigned int c[4] = {i, s, 0u, 0u};
  or_philox4x32_10(c, (unsigned int)(seed & 0xffffffffu),
                   (unsigned int)(seed >> 32));
  unsigned long long a = ((unsigned long long)c[0] << 32) | c[1];
  unsigned long long b = ((unsigned long long)c[2] << 32) | c[3];
  double u1 = ((double)(a >> 11) + 0.5) / 9007199254740992.0;
  double u2 = ((double)(b >> 11) + 0.5) / 9007199254740992.0;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ------------------------------------------------------------------ */
/* latent model: neighbors-of-neighbors index and the Gibbs sampler     */
/* ------------------------------------------------------------------ */

/* src/util.cpp:248-258 */
static int which_slot(int a, const int *b, int n) {
  for (int i = 0; i < n; i++)
    if (b[i] == a) return i;
  return -1;
}

/* src/nn.cpp:25-56.  searchType 0: the full scan of src/util.cpp:72-91
 * (children ascending); 1: the reversed scan of :93-112 (children
 * descending); otherwise the CSR -> CSC transpose of :115-185, whose row
 * pointers are the running sum of the counts (children ascending). */
int or_mkUIndx(int n, int m, const int *nnIndx, int *uIndx, int *uIndxLU,
               int *uiIndx, const int *nnIndxLU, int searchType) {
  if (check_nn_args(n, m)) return OR_ERR_ARG;
  const long long nIndx = n_indx(n, m);
  if (searchType == 0 || searchType == 1) {
    long long l = 0;
    for (int i = 0; i < n; i++) {
      uIndxLU[i] = (int)l;
      int h = 0;
      int j0 = searchType == 0 ? 0 : n - 1;
      int j1 = searchType == 0 ? n : i;
      int dj = searchType == 0 ? 1 : -1;
      for (int j = j0; j != j1; j += dj) {
        int off, cnt;
        or_getNNIndx(j, m, &off, &cnt);
        for (int k = 0; k < cnt; k++)
          if (nnIndx[off + k] == i) uIndx[l + h++] = j;
      }
      l += h;
      uIndxLU[n + i] = h;
    }
  } else {
    int *rp = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *cp = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    if (!rp || !cp) {
      free(rp);
      free(cp);
      return fail(OR_ERR_ARG, "allocation failed");
    }
    rp[0] = 0;
    for (int i = 1; i < n; i++) rp[i] = rp[i - 1] + nnIndxLU[n + i - 1];
    rp[n] = (int)nIndx;
    /* counts per column, exclusive running sum, then rows scattered in
     * ascending order (crs_csc) */
    for (int c = 0; c <= n; c++) cp[c] = 0;
    for (long long t = 0; t < nIndx; t++) cp[nnIndx[t]]++;
    int run = 0;
    for (int c = 0; c < n; c++) {
      int k = cp[c];
      cp[c] = run;
      run += k;
    }
    cp[n] = (int)nIndx;
    for (int r = 0; r < n; r++)
      for (int t = rp[r]; t < rp[r + 1]; t++) uIndx[cp[nnIndx[t]]++] = r;
    int last = 0;
    for (int c = 0; c <= n; c++) {
      int k = cp[c];
      cp[c] = last;
      last = k;
    }
    for (int i = 0; i < n; i++) {
      uIndxLU[i] = cp[i];
      uIndxLU[n + i] = cp[i + 1] - cp[i];
    }
    free(rp);
    free(cp);
  }
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < uIndxLU[n + i]; j++) {
      int k = uIndx[uIndxLU[i] + j];
      int s = which_slot(i, &nnIndx[nnIndxLU[k]], nnIndxLU[n + k]);
      if (s < 0) return fail(OR_ERR_ARG, "c++ Rf_error: which failed");
      uiIndx[uIndxLU[i] + j] = s;
    }
  }
  return OR_OK;
}

/* Standard exponential deviate: Ahrens & Dieter (1972), algorithm SA, as
 * R's exp_rand.  qtab[k] = sum_{j=1..k+1} (ln 2)^j / j!. */
static double r_exp_rand(const or_rng *r) {
  static const double qtab[16] = {
      0.6931471805599453, 0.933373687519046,  0.9888777961838676,
      0.9984959252914961, 0.9998292811061389, 0.9999833164100728,
      0.9999985691438767, 0.9999998906925558, 0.9999999924734159,
      0.9999999995283275, 0.9999999999728814, 0.9999999999985598,
      0.999999999999929,  0.9999999999999968, 0.9999999999999999, 1.0};
  double acc = 0.0;
  double u = r->unif_rand(r->state);
  while (u <= 0.0 || u >= 1.0) u = r->unif_rand(r->state);
  for (;;) {
    u += u;
    if (u > 1.0) break;
    acc += qtab[0];
  }
  u -= 1.0;
  if (u <= qtab[0]) return acc + u;
  int i = 0;
  double us = r->unif_rand(r->state), umin = us;
  do {
    us = r->unif_rand(r->state);
    if (us < umin) umin = us;
    i++;
  } while (u > qtab[i]);
  return acc + umin * qtab[0];
}

/* Gamma(a, scale) deviate for a >= 1: Ahrens & Dieter (1982), algorithm GD
 * (Comm. ACM 25:47-54), the method of R's rgamma for a >= 1, with the
 * hat-acceptance exp(q) - 1 taken as expm1(q).  Returns NaN for a < 1. */
static double gd_quot(double t, double s, double s2, double q0) {
  const double v = t / (s + s);
  if (fabs(v) <= 0.25) {
    double pa = 0.1233795;
    pa = pa * v - 0.1367177;
    pa = pa * v + 0.1423657;
    pa = pa * v - 0.1662921;
    pa = pa * v + 0.2000062;
    pa = pa * v - 0.250003;
    pa = pa * v + 0.3333333;
    return q0 + 0.5 * t * t * pa * v;
  }
  return q0 - s * t + 0.25 * t * t + (s2 + s2) * log(1.0 + v);
}

double or_rgamma(const or_rng *r, double a, double scale) {
  if (!(a >= 1.0) || !(scale > 0.0)) return NAN;
  const double s2 = a - 0.5, s = sqrt(s2);
  const double d = 5.656854 - 12.0 * s;
  /* immediate acceptance: t >= 0 */
  double t = r->norm_rand(r->state);
  double x = s + 0.5 * t;
  const double xx = x * x;
  if (t >= 0.0) return scale * xx;
  /* squeeze acceptance */
  double u = r->unif_rand(r->state);
  if (d * u <= t * t * t) return scale * xx;
  const double rr = 1.0 / a;
  double q0 = 2.424e-4;
  q0 = q0 * rr + 2.4511e-4;
  q0 = q0 * rr - 7.388e-5;
  q0 = q0 * rr + 0.00144121;
  q0 = q0 * rr + 0.00801191;
  q0 = q0 * rr + 0.02083148;
  q0 = q0 * rr + 0.04166669;
  q0 = q0 * rr;
  double b, si, c;
  if (a <= 3.686) {
    b = 0.463 + s + 0.178 * s2;
    si = 1.235;
    c = 0.195 / s - 0.079 + 0.16 * s;
  } else if (a <= 13.022) {
    b = 1.654 + 0.0076 * s2;
    si = 1.68 / s + 0.275;
    c = 0.062 / s + 0.024;
  } else {
    b = 1.77;
    si = 0.75;
    c = 0.1515 / s;
  }
  /* quotient acceptance */
  if (x > 0.0 && log(1.0 - u) <= gd_quot(t, s, s2, q0)) return scale * xx;
  /* double-exponential rejection */
  for (;;) {
    const double e = r_exp_rand(r);
    u = r->unif_rand(r->state);
    u = u + u - 1.0;
    t = u < 0.0 ? b - si * e : b + si * e;
    if (t < -0.71874483771719) continue;
    const double q = gd_quot(t, s, s2, q0);
    if (q <= 0.0) continue;
    if (c * fabs(u) <= expm1(q) * exp(e - 0.5 * t * t)) break;
  }
  x = s + 0.5 * t;
  return scale * x * x;
}

/* One site of the latent Gibbs sweep, src/sNNGP.cpp:260-287: the full
 * conditional of w_i given its neighbors, the sites that hold it as a
 * neighbor and their other neighbors; w_i <- mu var + sqrt(var) z. */
static void latent_site(int i, int n, int p, const double *y, const double *X,
                        const double *beta, double tau2, const double *B,
                        const double *F, const int *nnIndx,
                        const int *nnIndxLU, const int *uIndx,
                        const int *uIndxLU, const int *uiIndx, double z,
                        double *w) {
  double a = 0.0, v = 0.0;
  for (int j = 0; j < uIndxLU[n + i]; j++) {
    const int jj = uIndx[uIndxLU[i] + j];
    const int o = nnIndxLU[jj];
    double b = 0.0;
    for (int k = 0; k < nnIndxLU[n + jj]; k++) {
      const int kk = nnIndx[o + k];
      if (kk != i) b += B[o + k] * w[kk];
    }
    const double aij = w[jj] - b;
    const double bij = B[o + uiIndx[uIndxLU[i] + j]];
    a += bij * aij / F[jj];
    v += bij * bij / F[jj];
  }
  double e = 0.0;
  for (int j = 0; j < nnIndxLU[n + i]; j++)
    e += B[nnIndxLU[i] + j] * w[nnIndx[nnIndxLU[i] + j]];
  const double mu = (y[i] - ddot(p, &X[i], n, beta, 1)) / tau2 + e / F[i] + a;
  const double var = 1.0 / (1.0 / tau2 + 1.0 / F[i] + v);
  w[i] = mu * var + sqrt(var) * z;
}

/* sum_i (w_i - B_i . w_N(i))^2 / F_i (and sum log F_i), the loops at
 * src/sNNGP.cpp:319-333 */
static double latent_quad(const double *B, const double *F, const double *w,
                          int n, const int *nnIndx, const int *nnIndxLU,
                          double *logDet) {
  double a = 0.0, ld = 0.0;
  for (int i = 0; i < n; i++) {
    double b;
    if (nnIndxLU[n + i] > 0) {
      double e = 0.0;
      for (int j = 0; j < nnIndxLU[n + i]; j++)
        e += B[nnIndxLU[i] + j] * w[nnIndx[nnIndxLU[i] + j]];
      b = w[i] - e;
    } else {
      b = w[i];
    }
    a += b * b / F[i];
    ld += log(F[i]);
  }
  if (logDet) *logDet = ld;
  return a;
}

/* src/sNNGP.cpp:90-482: latent NNGP sampler.  Per iteration: the w sweep
 * (n normals, site order), beta | w (p normals), tau.sq | w, beta and
 * sigma.sq | w (one Gamma deviate each), then a Metropolis step on phi
 * [and nu] with B, F recomputed at the new sigma.sq (C = sigma.sq R, no
 * nugget; src/sNNGP.cpp:24-73).  visit == NULL sweeps the sites in order
 * 0..n-1 as the reference does; otherwise in the order visit[0..n-1],
 * with the same n normals drawn in site order before the sweep (site i
 * always uses normal i).  pWSamples n x nSamples. */
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
             double *pWSamples, int *nAccept) {
  if (covModel < 0 || covModel > 3)
    return fail(OR_ERR_COV_MODEL,
                "c++ Rf_error: cov.model is not correctly specified");
  const int matern = covModel == 2;
  const int nTheta = matern ? 4 : 3;
  const long long nIndx = n_indx(n, m);
  const size_t nb = (size_t)(nIndx > 0 ? nIndx : 1);
  double *B = (double *)malloc(sizeof(double) * nb);
  double *F = (double *)malloc(sizeof(double) * (size_t)n);
  double *Bc = (double *)malloc(sizeof(double) * nb);
  double *Fc = (double *)malloc(sizeof(double) * (size_t)n);
  double *w = (double *)calloc((size_t)n, sizeof(double));
  double *z = (double *)malloc(sizeof(double) * (size_t)n);
  double *tn = (double *)malloc(sizeof(double) * (size_t)n);
  double *XtX = (double *)malloc(sizeof(double) * (size_t)p * p + 1);
  double *tpp = (double *)malloc(sizeof(double) * (size_t)p * p + 1);
  double *tp = (double *)malloc(sizeof(double) * (size_t)p + 1);
  double *tp2 = (double *)malloc(sizeof(double) * (size_t)p + 1);
  double *beta = (double *)malloc(sizeof(double) * (size_t)p + 1);
  int st = OR_OK;
  if (!B || !F || !Bc || !Fc || !w || !z || !tn || !XtX || !tpp || !tp ||
      !tp2 || !beta) {
    st = fail(OR_ERR_ARG, "allocation failed");
    goto done;
  }
  {
    double th[4] = {sigmaSqStarting, tauSqStarting, phiStarting,
                    matern ? nuStarting : 0.0};
    const double sA = sigmaSqIG[0], sB = sigmaSqIG[1];
    const double tA = tauSqIG[0], tB = tauSqIG[1];
    const double pA = phiUnif[0], pB = phiUnif[1];
    const double nA = matern ? nuUnif[0] : 0.0, nB = matern ? nuUnif[1] : 0.0;
    for (int j = 0; j < p; j++) beta[j] = betaStarting[j];
    /* X^T X (dgemm 'T','N', src/sNNGP.cpp:244) */
    for (int j = 0; j < p; j++)
      for (int i = 0; i < p; i++) {
        double s = 0.0;
        for (int l = 0; l < n; l++) s += X[l + (size_t)i * n] * X[l + (size_t)j * n];
        XtX[i + j * p] = s;
      }
    st = bf_common(B, F, coords, nnIndx, nnIndxLU, n, m, th[0], 0.0, th[2],
                   th[3], covModel, nThreads);
    int accept = 0;
    for (int s = 0; s < nSamples && st == OR_OK; s++) {
      /* w | rest */
      for (int i = 0; i < n; i++) z[i] = r_rnorm(rng, 0.0, 1.0);
      for (int t = 0; t < n; t++) {
        const int i = visit ? visit[t] : t;
        latent_site(i, n, p, y, X, beta, th[1], B, F, nnIndx, nnIndxLU, uIndx,
                    uIndxLU, uiIndx, z[i], w);
      }
      /* beta | rest (src/sNNGP.cpp:292-308) */
      for (int i = 0; i < n; i++) tn[i] = (y[i] - w[i]) / th[1];
      for (int j = 0; j < p; j++) tp[j] = ddot(n, &X[(size_t)j * n], 1, tn, 1);
      for (int k = 0; k < p * p; k++) tpp[k] = XtX[k] / th[1];
      if (or_potrf_L(p, tpp) || or_potri_L(p, tpp)) {
        st = fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
        break;
      }
      or_symv_L(p, tpp, tp, tp2);
      if (or_potrf_L(p, tpp)) {
        st = fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
        break;
      }
      mvrnorm(rng, beta, tp2, tpp, p);
      /* tau.sq | rest (:313-317) */
      for (int i = 0; i < n; i++)
        tn[i] = y[i] - w[i] - ddot(p, &X[i], n, beta, 1);
      th[1] = 1.0 / or_rgamma(rng, tA + n / 2.0,
                              1.0 / (tB + 0.5 * ddot(n, tn, 1, tn, 1)));
      /* sigma.sq | rest (:322-341), B and F of the previous sigma.sq */
      double a = latent_quad(B, F, w, n, nnIndx, nnIndxLU, NULL);
      th[0] = 1.0 / or_rgamma(rng, sA + n / 2.0, 1.0 / (sB + 0.5 * a * th[0]));
      /* phi [, nu] | rest: Metropolis on the logit scale (:343-420) */
      st = bf_common(B, F, coords, nnIndx, nnIndxLU, n, m, th[0], 0.0, th[2],
                     th[3], covModel, nThreads);
      if (st) break;
      double ld;
      a = latent_quad(B, F, w, n, nnIndx, nnIndxLU, &ld);
      double lpCur = -0.5 * ld - 0.5 * a;
      lpCur += log(th[2] - pA) + log(pB - th[2]);
      if (matern) lpCur += log(th[3] - nA) + log(nB - th[3]);
      const double phiC =
          logit_inv(r_rnorm(rng, logit_ab(th[2], pA, pB), phiTuning), pA, pB);
      double nuC = 0.0;
      if (matern)
        nuC = logit_inv(r_rnorm(rng, logit_ab(th[3], nA, nB), nuTuning), nA,
                        nB);
      st = bf_common(Bc, Fc, coords, nnIndx, nnIndxLU, n, m, th[0], 0.0, phiC,
                     nuC, covModel, nThreads);
      if (st) break;
      a = latent_quad(Bc, Fc, w, n, nnIndx, nnIndxLU, &ld);
      double lpCand = -0.5 * ld - 0.5 * a;
      lpCand += log(phiC - pA) + log(pB - phiC);
      if (matern) lpCand += log(nuC - nA) + log(nB - nuC);
      if (r_runif(rng, 0.0, 1.0) <= exp(lpCand - lpCur)) {
        double *tb = B;
        B = Bc;
        Bc = tb;
        double *tf = F;
        F = Fc;
        Fc = tf;
        th[2] = phiC;
        if (matern) th[3] = nuC;
        accept++;
      }
      for (int j = 0; j < p; j++) pBetaSamples[(size_t)s * p + j] = beta[j];
      for (int k = 0; k < nTheta; k++)
        pThetaSamples[(size_t)s * nTheta + k] = th[k];
      for (int i = 0; i < n; i++) pWSamples[(size_t)s * n + i] = w[i];
    }
    if (nAccept) *nAccept = accept;
  }
done:
  free(B);
  free(F);
  free(Bc);
  free(Fc);
  free(w);
  free(z);
  free(tn);
  free(XtX);
  free(tpp);
  free(tp);
  free(tp2);
  free(beta);
  return st;
}

/* src/sNNGPPredict.cpp:25-241, Gaussian family, in the single-thread order:
 * normal index i*S + s for site i and sample s, w0's normals first, then
 * y0's (z holds 2 q S).  Without a nugget: C = s2 R(N,N), c = s2 R(N,i). */
int or_sNNGPPredict(const double *X0, const double *coords, int n, int p,
                    int m, const double *coords0, int q, const int *nnIndx0,
                    const double *beta, const double *theta, const double *w,
                    int nSamples, int covModel, const double *z, double *y0,
                    double *w0) {
  if (covModel < 0 || covModel > 3)
    return fail(OR_ERR_COV_MODEL,
                "c++ Rf_error: cov.model is not correctly specified");
  const int nTheta = covModel == 2 ? 4 : 3;
  double *C = (double *)malloc(sizeof(double) * (size_t)m * m);
  double *c = (double *)malloc(sizeof(double) * (size_t)m);
  double *u = (double *)malloc(sizeof(double) * (size_t)m);
  if (!C || !c || !u) {
    free(C);
    free(c);
    free(u);
    return fail(OR_ERR_ARG, "allocation failed");
  }
  const size_t qs = (size_t)q * nSamples;
  int st = OR_OK;
  for (int i = 0; i < q && st == OR_OK; i++) {
    for (int s = 0; s < nSamples; s++) {
      const double *th = theta + (size_t)s * nTheta;
      const double s2 = th[0], t2 = th[1], phi = th[2];
      const double nu = covModel == 2 ? th[3] : 0.0;
      for (int k = 0; k < m; k++) {
        const int a = nnIndx0[i + (size_t)q * k];
        double d = or_dist2(coords[a], coords[n + a], coords0[i],
                            coords0[q + i]);
        c[k] = s2 * or_spCor(d, phi, nu, covModel);
        for (int l = 0; l < m; l++) {
          const int b = nnIndx0[i + (size_t)q * l];
          d = or_dist2(coords[a], coords[n + a], coords[b], coords[n + b]);
          C[l * m + k] = s2 * or_spCor(d, phi, nu, covModel);
        }
      }
      if (or_potrf_L(m, C) || or_potri_L(m, C)) {
        st = fail(OR_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
        break;
      }
      or_symv_L(m, C, c, u);
      double d = 0.0;
      for (int k = 0; k < m; k++)
        d += u[k] * w[(size_t)s * n + nnIndx0[i + (size_t)q * k]];
      const size_t zi = (size_t)i * nSamples + s;
      const double wv = sqrt(s2 - ddot(m, u, 1, c, 1)) * z[zi] + d;
      w0[(size_t)s * q + i] = wv;
      y0[(size_t)s * q + i] = sqrt(t2) * z[qs + zi] +
                              ddot(p, &X0[i], q, &beta[(size_t)s * p], 1) + wv;
    }
  }
  free(C);
  free(c);
  free(u);
  return st;
}
