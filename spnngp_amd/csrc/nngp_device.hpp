// nngp_device.hpp — device-side math shared by the NNGP kernels (gfx950).
//
// Covariance functions follow src/util.cpp:208-245 (spCor) of spNNGP 1.0.1:
//   0 exponential  exp(-phi D)
//   1 spherical    1 - 1.5 phi D + 0.5 (phi D)^3 for 0 < D <= 1/phi, 0 beyond
//   2 matern       (D phi)^nu / (2^(nu-1) Gamma(nu)) K_nu(D phi)
//   3 gaussian     exp(-(phi D)^2)
// every one scaled by s2 (sigma.sq; 1 on the correlation scale of cNNGP).
// The host fills CovArgs once per theta, so per-pair work is only the
// distance, the special function and the scale.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nngp {

constexpr int kExpDeg = 11;  // Taylor degree of exp on |r| <= ln2/2
constexpr int kExp2Deg = 10; // degree of the 2^f polynomial on |f| <= 1/2

struct CovArgs {
  double s2;       // sigma.sq (1 on the correlation scale)
  double t2;       // tau.sq (alpha on the correlation scale)
  double phi;
  double nu;
  double diag;     // s2 * rho(0) + t2  (rho(0) == 1 for every model)
  double inv_phi;  // 1.0 / phi, the spherical support (src/util.cpp:221)
  double phi2;     // phi * phi (gaussian on squared distances)
  double log2e;    // 1 / ln 2
  double ln2_hi;   // ln 2 split so that k * ln2_hi is exact for |k| < 2^20
  double ln2_lo;
  double ec[kExpDeg + 1];  // s2 / k!: exp polynomial with the scale folded in
  double ec2[kExpDeg + 1]; // s2 times the coefficients of the degree-10
                           // polynomial of 2^f on |f| <= 1/2 (Chebyshev-node
                           // interpolant); entry 11 unused
  double tsc;      // coordinate scale of the per-lane updateBF kernels:
                   // phi / ln 2 (exponential), phi / sqrt(ln 2) (gaussian)
  // matern (Temme's method; see matern_s2 below)
  double mnorm;    // s2 / (2^(nu-1) Gamma(nu))
  double mu;       // nu - nl, in [-1/2, 1/2)
  double g1, g2;   // Gamma_1(mu), Gamma_2(mu)
  double rgp;      // 1 / Gamma(1 + mu)
  double rgm;      // 1 / Gamma(1 - mu)
  double gp, gm;   // Gamma(1 + mu), Gamma(1 - mu)
  double fa;       // pi mu / sin(pi mu) (1 at mu = 0)
  int nl;          // round(nu)
  int cov;         // covModel
};

// ---------------------------------------------------------------------------
// elementary functions
// ---------------------------------------------------------------------------

// sqrt(s) for s >= 0: hardware reciprocal square root and one Newton step
// (relative error below 1e-14; exact 0 for s == 0).
__device__ __forceinline__ double nn_sqrt(double s) {
  double y = __builtin_amdgcn_rsq(fmax(s, 1e-300));
  double d = s * y;
  double r = fma(-d, d, s);
  return fma(r * y, 0.5, d);
}

// 1/sqrt(a) for a > 0 to full double precision (two Newton steps).
__device__ __forceinline__ double nn_rsqrt(double a) {
  double y = __builtin_amdgcn_rsq(a);
  double e = fma(-(a * y), y, 1.0);
  y = fma(y * e, 0.5, y);
  e = fma(-(a * y), y, 1.0);
  return fma(y * e, 0.5, y);
}

// s2 * exp(x) for x <= 0 (large negative x gives 0).  Range reduction
// x = k ln2 + r, |r| <= ln2/2, Taylor polynomial of degree 11 whose
// coefficients s2/j! live in SGPRs (relative error ~1e-15).
__device__ __forceinline__ double s2_exp(double x, const CovArgs &a) {
  x = fmax(x, -1000.0);
  double k = __builtin_rint(x * a.log2e);
  double r = fma(k, -a.ln2_hi, x);
  r = fma(k, -a.ln2_lo, r);
  double p = a.ec[11];
#pragma unroll
  for (int j = 10; j >= 0; j--) p = fma(p, r, a.ec[j]);
  return ldexp(p, (int)k);
}

// s2 * 2^t for t <= 0 (large negative t gives 0): t = k + f with k the
// nearest integer, so |f| <= 1/2 and f = t - k is exact, then a degree-10
// polynomial of 2^f, the interpolant at the Chebyshev nodes of [-1/2, 1/2]
// with s2 folded into its coefficients (relative error 4.5e-16 before
// rounding).  With t = -phi d / ln 2 (or -phi^2 d^2 / ln 2) formed from
// coordinates scaled by tsc, this is s2_corr of the exponential (gaussian)
// covariance without the products by phi and 1 / ln 2, without the
// two-step reduction, and with one multiply-add fewer than s2_exp.
__device__ __forceinline__ double s2_exp2(double t, const CovArgs &a) {
  t = fmax(t, -1500.0);
  const double k = __builtin_rint(t);
  const double f = t - k;
  double p = a.ec2[kExp2Deg];
#pragma unroll
  for (int j = kExp2Deg - 1; j >= 0; j--) p = fma(p, f, a.ec2[j]);
  return ldexp(p, (int)k);
}

// s2_exp2 for finite t <= 0, without the clamp: for t far below -1075 the
// conversion of k to int saturates at INT_MIN and ldexp returns 0, and f =
// t - k stays exact (0 once |t| >= 2^52), so the result is the same.  Only
// an infinite t (which the per-lane kernels never form: their padding sites
// sit at finite coordinates) would give NaN here.  t must arrive rounded:
// the per-lane kernels pass -sqrt(d2) or -d2 from a fused multiply-add.  A
// product passed as t could be contracted into t - k, which would leave f
// as the product's rounding error (about 1e84 for a distance of 1e100)
// instead of 0; the clamp of s2_exp2 prevents that contraction.
__device__ __forceinline__ double s2_exp2_finite(double t,
                                                 const CovArgs &a) {
  const double k = __builtin_rint(t);
  const double f = t - k;
  double p = a.ec2[kExp2Deg];
#pragma unroll
  for (int j = kExp2Deg - 1; j >= 0; j--) p = fma(p, f, a.ec2[j]);
  return ldexp(p, (int)k);
}

// K_nu(x) for x > 0 by Temme's series (x <= 2) or Steed's continued
// fraction (x > 2) for K_mu, K_{mu+1}, |mu| <= 1/2, then upward recurrence.
__device__ inline double bessel_k(double x, const CovArgs &a) {
  const double eps = 1e-17;
  const double pi = 3.14159265358979323846;
  const double mu = a.mu;
  double kmu, kmu1;
  if (x <= 2.0) {
    double half = 0.5 * x;
    double lg = -log(half);
    double sg = mu * lg;
    double pm = pi * mu;
    double fa = fabs(pm) < 1e-300 ? 1.0 : pm / sin(pm);
    double fs = fabs(sg) < 1e-300 ? 1.0 : sinh(sg) / sg;
    double f = fa * (a.g1 * cosh(sg) + a.g2 * fs * lg);
    double es = exp(sg);
    double pk = 0.5 * es / a.rgp;
    double qk = 0.5 / (es * a.rgm);
    double ck = 1.0, d = half * half;
    double s0 = f, s1 = pk;
    for (int k = 1; k < 200; k++) {
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
    double b = 2.0 * (1.0 + x);
    double d = 1.0 / b;
    double h = d, dh = d;
    double q1 = 0.0, q2 = 1.0;
    double a1 = 0.25 - mu * mu;
    double q = a1, c = a1;
    double aa = -a1;
    double s = 1.0 + q * dh;
    for (int i = 2; i < 2000; i++) {
      aa -= 2.0 * (i - 1);
      c = -aa * c / i;
      double qn = (q1 - b * q2) / aa;
      q1 = q2;
      q2 = qn;
      q += c * qn;
      b += 2.0;
      d = 1.0 / (b + aa * d);
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
  for (int i = 1; i <= a.nl; i++) {
    double kn = (mu + i) * (2.0 / x) * kmu1 + kmu;
    kmu = kmu1;
    kmu1 = kn;
  }
  return kmu;
}

// Reciprocals the Bessel-K loops divide by, which depend only on the loop
// index and mu: filled once per block (bessel_tab_fill) so the series and
// continued fraction multiply instead of divide.
constexpr int kBesTab = 128;
struct BesselTab {
  double ik[kBesTab];    // 1 / k
  double ik2[kBesTab];   // 1 / (k^2 - mu^2)
  double ikm[kBesTab];   // 1 / (k - mu)
  double ikp[kBesTab];   // 1 / (k + mu)
  double iaa[kBesTab];   // 1 / (mu^2 - 1/4 - k (k - 1)), Steed's a_k
};

// Every thread of the block calls it (then a barrier before use).
__device__ inline void bessel_tab_fill(BesselTab &t, const CovArgs &a) {
  const double mu = a.mu;
  for (int k = (int)threadIdx.x; k < kBesTab; k += (int)blockDim.x) {
    const double dk = (double)(k > 0 ? k : 1);
    t.ik[k] = 1.0 / dk;
    t.ik2[k] = 1.0 / (dk * dk - mu * mu);
    t.ikm[k] = 1.0 / (dk - mu);
    t.ikp[k] = 1.0 / (dk + mu);
    t.iaa[k] = 1.0 / (mu * mu - 0.25 - dk * (dk - 1.0));
  }
}

// 1/y to ~1 ulp: hardware reciprocal and two Newton steps.
__device__ __forceinline__ double nn_rcp(double y) {
  double r = __builtin_amdgcn_rcp(y);
  double e = fma(-y, r, 1.0);
  r = fma(r, e, r);
  e = fma(-y, r, 1.0);
  return fma(r, e, r);
}

// x^nu K_nu(x) for 0 < x <= 700: the method of bessel_k above (Temme's
// series for x <= 2, Steed's continued fraction above, upward recurrence in
// nu) with the loop divisors taken from t, the powers of x folded into one
// exp, and sinh / cosh of the series built from one exp.
__device__ inline double xnu_bessel_k(double x, const CovArgs &a,
                                      const BesselTab &t) {
  const double eps = 1e-17;
  const double mu = a.mu;
  const double tx = nn_rcp(x) * 2.0;  // 2 / x
  double kmu, kmu1, scale;
  if (x <= 2.0) {
    const double lg = -log(0.5 * x);
    const double sg = mu * lg;
    const double es = exp(sg);
    const double ies = nn_rcp(es);
    double fs;  // sinh(sg) / sg
    if (fabs(sg) < 0.1) {
      const double s2 = sg * sg;
      fs = 1.0 + s2 * (1.0 / 6.0 + s2 * (1.0 / 120.0 + s2 * (1.0 / 5040.0 +
           s2 * (1.0 / 362880.0 + s2 * (1.0 / 39916800.0)))));
    } else {
      fs = 0.5 * (es - ies) / sg;
    }
    double f = a.fa * (a.g1 * (0.5 * (es + ies)) + a.g2 * fs * lg);
    double pk = 0.5 * es * a.gp;
    double qk = 0.5 * ies * a.gm;
    double ck = 1.0;
    const double d = 0.25 * x * x;
    double s0 = f, s1 = pk;
    for (int k = 1; k < 200; k++) {
      double i2, i1, im, ip;
      if (k < kBesTab) {
        i2 = t.ik2[k];
        i1 = t.ik[k];
        im = t.ikm[k];
        ip = t.ikp[k];
      } else {
        i2 = 1.0 / ((double)k * k - mu * mu);
        i1 = 1.0 / k;
        im = 1.0 / (k - mu);
        ip = 1.0 / (k + mu);
      }
      f = (k * f + pk + qk) * i2;
      ck *= d * i1;
      pk *= im;
      qk *= ip;
      const double t0 = ck * f;
      s0 += t0;
      s1 += ck * (pk - k * f);
      if (fabs(t0) < fabs(s0) * eps) break;
    }
    kmu = s0;
    kmu1 = s1 * tx;
    scale = exp(a.nu * (0.6931471805599453 - lg));  // x^nu
  } else {
    double b = 2.0 * (1.0 + x);
    double dd = nn_rcp(b);
    double h = dd, dh = dd;
    double q1 = 0.0, q2 = 1.0;
    const double a1 = 0.25 - mu * mu;
    double q = a1, c = a1;
    double aa = -a1;
    double s = 1.0 + q * dh;
    for (int i = 2; i < 2000; i++) {
      aa -= 2.0 * (i - 1);
      double inv_i, inv_aa;
      if (i < kBesTab) {
        inv_i = t.ik[i];
        inv_aa = t.iaa[i];
      } else {
        inv_i = 1.0 / i;
        inv_aa = 1.0 / aa;
      }
      c = -aa * c * inv_i;
      const double qn = (q1 - b * q2) * inv_aa;
      q1 = q2;
      q2 = qn;
      q += c * qn;
      b += 2.0;
      dd = nn_rcp(b + aa * dd);
      dh = (b * dd - 1.0) * dh;
      h += dh;
      const double ds = q * dh;
      s += ds;
      if (fabs(ds) < fabs(s) * eps) break;
    }
    h = a1 * h;
    // K_mu without its e^-x: sqrt(pi / (2x)) / s; e^-x x^nu folded below
    kmu = 1.2533141373155003 * nn_rsqrt(x) * nn_rcp(s);
    kmu1 = kmu * (mu + x + 0.5 - h) * (0.5 * tx);
    scale = exp(a.nu * log(x) - x);
  }
  for (int i = 1; i <= a.nl; i++) {
    const double kn = fma((mu + i) * tx, kmu1, kmu);
    kmu = kmu1;
    kmu1 = kn;
  }
  return kmu * scale;
}

// s2 * matern correlation at distance d (src/util.cpp:228-237).
__device__ inline double matern_s2(double d, const CovArgs &a) {
  double x = d * a.phi;
  if (!(x > 0.0)) return a.s2;
  if (x > 700.0) return 0.0;  // K_nu(x) x^nu underflows long before this
  return a.mnorm * pow(x, a.nu) * bessel_k(x, a);
}

// The same with the loop divisors from a block's table.
__device__ inline double matern_s2_tab(double d, const CovArgs &a,
                                       const BesselTab &t) {
  const double x = d * a.phi;
  if (!(x > 0.0)) return a.s2;
  if (x > 700.0) return 0.0;
  return a.mnorm * xnu_bessel_k(x, a, t);
}

// ---------------------------------------------------------------------------
// the general-nu matern as a table per pass
// ---------------------------------------------------------------------------
// s2 * rho(x) = mnorm * x^nu K_nu(x) for x in [2^-30, 2^6) as polynomials,
// one per half octave: x = m 2^e with m in [1, sqrt 2) or [sqrt 2, 2), and
// on each piece a degree-15 polynomial in u = (m - c) / h in [-1, 1], the
// interpolant at the 16 Chebyshev points (built by matern_tab_fill from
// xnu_bessel_k).  x^nu K_nu(x) = A(x^2) + x^(2 nu) B(x^2) with A, B entire,
// so its only singularity is x = 0, three half-widths from the nearest
// piece's centre on this grid; the interpolants converge like 11.6^-n and
// agree with the series to ~1e-14 relative (~5e-15 absolute).  Outside the
// range (near-coincident sites, or x >= 64 where the entry is below 1e-25)
// the series itself is used.
constexpr int kMatOct0 = -30;             // first octave 2^-30
constexpr int kMatOct = 36;               // octaves, up to 2^6
constexpr int kMatDeg = 15;
constexpr int kMatN = kMatDeg + 1;        // nodes and coefficients a piece
constexpr int kMatPieces = 2 * kMatOct;
constexpr double kSqrt2 = 1.4142135623730951;

// Rows padded by one double: lanes evaluating neighboring pieces then read
// their k-th coefficients from different LDS banks (at 16 doubles a row,
// every piece's k-th coefficient sat in the same bank).
struct MaternTab {
  double a[kMatPieces][kMatN + 1];  // monomial coefficients in u, s2 folded
};

// Piece of mantissa m in [1, 2): lo [1, sqrt 2), hi [sqrt 2, 2); u = m / h
// - c / h with centre c and half-width h.
__host__ __device__ constexpr double mat_c(int hi) {
  return hi ? (kSqrt2 + 2.0) * 0.5 : (1.0 + kSqrt2) * 0.5;
}
__host__ __device__ constexpr double mat_h(int hi) {
  return hi ? (2.0 - kSqrt2) * 0.5 : (kSqrt2 - 1.0) * 0.5;
}

// Monomial coefficients of the Chebyshev polynomials T_0 .. T_15 (exact
// integers): T_{k+1} = 2u T_k - T_{k-1}.
struct ChebMonomials {
  double t[kMatN][kMatN];
  constexpr ChebMonomials() : t() {
    t[0][0] = 1.0;
    t[1][1] = 1.0;
    for (int k = 2; k < kMatN; k++)
      for (int i = 0; i < kMatN; i++)
        t[k][i] = (i > 0 ? 2.0 * t[k - 1][i - 1] : 0.0) - t[k - 2][i];
  }
};
__device__ constexpr ChebMonomials kChebMono{};

// Fills T for the pass's matern (every thread of the block calls it, after
// the BesselTab bt is filled and visible).  scratch: 2 kMatPieces kMatN +
// kMatN^2 doubles of LDS that are free until the call returns.
__device__ inline void matern_tab_fill(MaternTab &T, const CovArgs &a,
                                       const BesselTab &bt, double *scratch) {
  constexpr double kPi = 3.14159265358979323846;
  double *fv = scratch;                             // node values
  double *cc = scratch + kMatPieces * kMatN;        // Chebyshev coefficients
  double *cs = cc + kMatPieces * kMatN;             // cos(k theta_j)
  const int nt = (int)blockDim.x, t0 = (int)threadIdx.x;
  for (int q = t0; q < kMatN * kMatN; q += nt) {
    const int k = q / kMatN, j = q % kMatN;
    cs[q] = cos(kPi * k * (j + 0.5) / kMatN);
  }
  for (int q = t0; q < kMatPieces * kMatN; q += nt) {
    const int pc = q / kMatN, j = q % kMatN, hi = pc & 1;
    const double u = cos(kPi * (j + 0.5) / kMatN);
    const double x = ldexp(mat_c(hi) + mat_h(hi) * u, kMatOct0 + pc / 2);
    fv[q] = a.mnorm * xnu_bessel_k(x, a, bt);
  }
  __syncthreads();
  for (int q = t0; q < kMatPieces * kMatN; q += nt) {
    const int pc = q / kMatN, k = q % kMatN;
    double s = 0.0;
    for (int j = 0; j < kMatN; j++) s = fma(fv[pc * kMatN + j], cs[k * kMatN + j], s);
    cc[q] = s * (k == 0 ? 1.0 / kMatN : 2.0 / kMatN);
  }
  __syncthreads();
  for (int q = t0; q < kMatPieces * kMatN; q += nt) {
    const int pc = q / kMatN, i = q % kMatN;
    double s = 0.0;
    for (int k = i; k < kMatN; k++) s = fma(cc[pc * kMatN + k], kChebMono.t[k][i], s);
    T.a[pc][i] = s;
  }
  __syncthreads();
}

// s2 * matern correlation at distance d from the table (the series outside
// its range).
// The table's octave of x > 0: inside the table when 0 <= o < kMatOct.
__device__ __forceinline__ int matern_tab_octave(double x) {
  return __builtin_amdgcn_frexp_exp(x) - 1 - kMatOct0;
}

// The table's s2 * rho at x, for x > 0 in octave o of the table.
__device__ __forceinline__ double matern_tab_at(double x, int o,
                                                const MaternTab &T) {
  const double m = 2.0 * __builtin_amdgcn_frexp_mant(x);  // in [1, 2)
  const int hi = m >= kSqrt2 ? 1 : 0;
  const double u = hi ? fma(m, 1.0 / mat_h(1), -mat_c(1) / mat_h(1))
                      : fma(m, 1.0 / mat_h(0), -mat_c(0) / mat_h(0));
  const double *p = T.a[2 * o + hi];
  double r = p[kMatDeg];
#pragma unroll
  for (int k = kMatDeg - 1; k >= 0; k--) r = fma(r, u, p[k]);
  return r;
}

__device__ __forceinline__ double matern_s2_fast(double d, const CovArgs &a,
                                                 const BesselTab &bt,
                                                 const MaternTab &T) {
  const double x = d * a.phi;
  if (!(x > 0.0)) return a.s2;
  const int o = matern_tab_octave(x);
  if ((unsigned)o >= (unsigned)kMatOct) {
    if (x > 700.0) return 0.0;
    return a.mnorm * xnu_bessel_k(x, a, bt);
  }
  return matern_tab_at(x, o, T);
}

// Covariance variants of the templated kernels: the four covModels, plus
// the matern at nu = 1.5 and nu = 2.5 in closed form, (1 + x) e^-x and
// (1 + x + x^2/3) e^-x with x = phi D (the Bessel-K expression of
// src/util.cpp:228-237 equals these identically; nu = 0.5 is exp(-x)).
constexpr int kCovMatern15 = 5;
constexpr int kCovMatern25 = 6;

// The templated variant for covModel cov at smoothness nu, or -1 when only
// the runtime (Bessel-K) evaluation applies.
__host__ __device__ inline int cov_variant(int cov, double nu) {
  if (cov != 2) return cov;
  if (nu == 0.5) return 0;
  if (nu == 1.5) return kCovMatern15;
  if (nu == 2.5) return kCovMatern25;
  return -1;
}

// s2 * rho(D) from the squared distance d2 (COV is a variant above).
template <int COV>
__device__ __forceinline__ double s2_corr(double d2, const CovArgs &a) {
  if constexpr (COV == 0) {
    return s2_exp(-a.phi * nn_sqrt(d2), a);
  } else if constexpr (COV == 1) {
    double D = nn_sqrt(d2);
    double t = a.phi * D;
    double v = a.s2 * (1.0 - 1.5 * t + 0.5 * (t * t * t));
    return (D > 0.0 && D <= a.inv_phi) ? v : (D >= a.inv_phi ? 0.0 : a.s2);
  } else if constexpr (COV == 2) {
    return matern_s2(nn_sqrt(d2), a);
  } else if constexpr (COV == kCovMatern15) {
    const double x = nn_sqrt(d2) * a.phi;
    return s2_exp(-x, a) * (1.0 + x);
  } else if constexpr (COV == kCovMatern25) {
    const double x = nn_sqrt(d2) * a.phi;
    return s2_exp(-x, a) * (1.0 + x + x * x * (1.0 / 3.0));
  } else {
    return s2_exp(-a.phi2 * d2, a);
  }
}

// The per-lane kernels build the exponential, gaussian and closed-form
// matern covariances from neighbor offsets to the site scaled by tsc, so
// that the exponential factor is 2^-d (2^-d^2) for s2_exp2.
template <int COV>
constexpr bool lane_base2() {
  return COV == 0 || COV == 3 || COV == kCovMatern15 || COV == kCovMatern25;
}

// Covariances of the per-lane kernels that take the square root of the
// squared distance: the exponential and the closed-form materns.
template <int COV>
constexpr bool lane_sqrt() {
  return COV == 0 || COV == kCovMatern15 || COV == kCovMatern25;
}

// Squared distance of an offset (dx, dy) for lane_corr.  Where lane_corr
// takes the square root, the floor 1e-300 that keeps rsq finite at zero
// distance (duplicate sites) enters as the addend of the first multiply-add
// instead of as a max before rsq: one instruction fewer on each entry's
// dependency chain.  The floor changes no squared distance above 1e-284,
// and a zero distance still gives the factor 2^-1e-150, which rounds to the
// factor at 0.
template <int COV>
__device__ __forceinline__ double lane_d2(double dx, double dy) {
  if constexpr (lane_sqrt<COV>()) return fma(dx, dx, fma(dy, dy, 1e-300));
  else return fma(dx, dx, dy * dy);
}

// nn_sqrt for s >= 1e-300, as lane_d2 gives it: no floor of its own.
__device__ __forceinline__ double nn_sqrt_floored(double s) {
  double y = __builtin_amdgcn_rsq(s);
  double d = s * y;
  double r = fma(-d, d, s);
  return fma(r * y, 0.5, d);
}

// Covariance entry of the per-lane kernels from a squared distance from
// lane_d2 (of offsets scaled by tsc where lane_base2<COV>(), as given
// otherwise).
template <int COV>
__device__ __forceinline__ double lane_corr(double d2, const CovArgs &c) {
  constexpr double kLn2 = 0.6931471805599453;  // phi d = d' ln 2
  if constexpr (COV == 0) {
    return s2_exp2_finite(-nn_sqrt_floored(d2), c);
  } else if constexpr (COV == 3) {
    return s2_exp2_finite(-d2, c);
  } else if constexpr (COV == kCovMatern15) {
    const double dd = nn_sqrt_floored(d2);
    return s2_exp2_finite(-dd, c) * (1.0 + dd * kLn2);
  } else if constexpr (COV == kCovMatern25) {
    const double dd = nn_sqrt_floored(d2);
    const double x = dd * kLn2;
    return s2_exp2_finite(-dd, c) * (1.0 + x + x * x * (1.0 / 3.0));
  } else {
    return s2_corr<COV>(d2, c);
  }
}

// Runtime-dispatched variant for the generic (non-templated) kernels.
__device__ inline double s2_corr_rt(double d2, const CovArgs &a) {
  switch (a.cov) {
    case 0: return s2_corr<0>(d2, a);
    case 1: return s2_corr<1>(d2, a);
    case 2: return s2_corr<2>(d2, a);
    default: return s2_corr<3>(d2, a);
  }
}

// The same, with the matern's Bessel-K loop divisors from a block's table.
__device__ inline double s2_corr_rt(double d2, const CovArgs &a,
                                    const BesselTab &t) {
  if (a.cov == 2) return matern_s2_tab(nn_sqrt(d2), a, t);
  return s2_corr_rt(d2, a);
}

// The same, with the matern from the block's table of the pass.
__device__ inline double s2_corr_rt(double d2, const CovArgs &a,
                                    const BesselTab &t, const MaternTab &mt) {
  if (a.cov == 2) return matern_s2_fast(nn_sqrt(d2), a, t, mt);
  return s2_corr_rt(d2, a);
}

// Squared distance of two points.
__device__ __forceinline__ double sqdist(double ax, double ay, double bx,
                                         double by) {
  double dx = ax - bx, dy = ay - by;
  return fma(dx, dx, dy * dy);
}

// Packed lower triangle index, a[k(k+1)/2 + l] for l <= k.
__host__ __device__ constexpr int tri(int k, int l) { return k * (k + 1) / 2 + l; }

// In-register solve of one M x M SPD system held as a packed lower triangle
// `a` (overwritten by the Cholesky factor, diagonal as 1/L_jj) with
// right-hand side `v` (overwritten by C^-1 v).  vv = v^T C^-1 v.  Returns
// false when a pivot is not positive (the reference's dpotrf failure).
template <int M>
__device__ __forceinline__ bool chol_solve(double (&a)[M * (M + 1) / 2],
                                           double (&v)[M], double &vv) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < M; j++) {
    const double piv = a[tri(j, j)];
    ok = ok && (piv > 0.0);
    const double r = nn_rsqrt(piv);
    a[tri(j, j)] = r;
#pragma unroll
    for (int k = j + 1; k < M; k++) a[tri(k, j)] *= r;
#pragma unroll
    for (int k = j + 1; k < M; k++) {
      const double lkj = a[tri(k, j)];
#pragma unroll
      for (int l = j + 1; l <= k; l++)
        a[tri(k, l)] = fma(-lkj, a[tri(l, j)], a[tri(k, l)]);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < M; k++) {
    double t = v[k];
#pragma unroll
    for (int l = 0; l < k; l++) t = fma(-a[tri(k, l)], v[l], t);
    t *= a[tri(k, k)];
    v[k] = t;
    s = fma(t, t, s);
  }
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {
    const double t = v[k] * a[tri(k, k)];
    v[k] = t;
#pragma unroll
    for (int l = 0; l < k; l++) v[l] = fma(-a[tri(k, l)], t, v[l]);
  }
  vv = s;
  return ok;
}

// ---------------------------------------------------------------------------
// FP32 pieces of the mixed-precision updateBF
// ---------------------------------------------------------------------------
// Covariance parameters rounded to FP32 once per launch.
struct CovArgs32 {
  float s2, phi, inv_phi, phi2, diag;
  // general matern (cov 2), as in CovArgs
  float mnorm, mu, g1, g2, gp, gm, fa;
  int nl, cov;
};

__device__ __forceinline__ CovArgs32 cov32(const CovArgs &a) {
  CovArgs32 c;
  c.s2 = (float)a.s2;
  c.phi = (float)a.phi;
  c.inv_phi = (float)a.inv_phi;
  c.phi2 = (float)a.phi2;
  c.diag = (float)a.diag;
  c.mnorm = (float)a.mnorm;
  c.mu = (float)a.mu;
  c.g1 = (float)a.g1;
  c.g2 = (float)a.g2;
  c.gp = (float)a.gp;
  c.gm = (float)a.gm;
  c.fa = (float)a.fa;
  c.nl = a.nl;
  c.cov = a.cov;
  return c;
}

// s2 * rho in FP32 from an FP32 squared distance (hardware exp2 and sqrt,
// about 1 ulp each).
template <int COV>
__device__ __forceinline__ float s2_corr32(float d2, const CovArgs32 &a) {
  constexpr float kLog2e = 1.4426950408889634f;
  if constexpr (COV == 0) {
    return a.s2 * __builtin_amdgcn_exp2f(-a.phi * __builtin_amdgcn_sqrtf(d2) * kLog2e);
  } else if constexpr (COV == 1) {
    const float D = __builtin_amdgcn_sqrtf(d2);
    const float t = a.phi * D;
    const float v = a.s2 * (1.0f - 1.5f * t + 0.5f * (t * t * t));
    return (D > 0.0f && D <= a.inv_phi) ? v : (D >= a.inv_phi ? 0.0f : a.s2);
  } else if constexpr (COV == kCovMatern15) {
    const float x = __builtin_amdgcn_sqrtf(d2) * a.phi;
    if (!(x < 100.0f)) return 0.0f;  // e^-x underflows FP32; keeps inf out
    return a.s2 * __builtin_amdgcn_exp2f(-x * kLog2e) * (1.0f + x);
  } else if constexpr (COV == kCovMatern25) {
    const float x = __builtin_amdgcn_sqrtf(d2) * a.phi;
    if (!(x < 100.0f)) return 0.0f;
    return a.s2 * __builtin_amdgcn_exp2f(-x * kLog2e) *
           (1.0f + x + x * x * (1.0f / 3.0f));
  } else {
    return a.s2 * __builtin_amdgcn_exp2f(-a.phi2 * d2 * kLog2e);
  }
}

// The Bessel-K loop divisors of BesselTab in FP32 (each rounded once from
// its FP64 value).
constexpr int kBesTab32 = 64;
struct BesselTab32 {
  float ik[kBesTab32], ik2[kBesTab32], ikm[kBesTab32], ikp[kBesTab32],
      iaa[kBesTab32];
};

// Every thread of the block calls it (then a barrier before use).
__device__ inline void bessel_tab_fill32(BesselTab32 &t, const CovArgs &a) {
  const double mu = a.mu;
  for (int k = (int)threadIdx.x; k < kBesTab32; k += (int)blockDim.x) {
    const double dk = (double)(k > 0 ? k : 1);
    t.ik[k] = (float)(1.0 / dk);
    t.ik2[k] = (float)(1.0 / (dk * dk - mu * mu));
    t.ikm[k] = (float)(1.0 / (dk - mu));
    t.ikp[k] = (float)(1.0 / (dk + mu));
    t.iaa[k] = (float)(1.0 / (mu * mu - 0.25 - dk * (dk - 1.0)));
  }
}

// x^nu K_nu(x) in FP32 for 0 < x < 80 (the mixed form's covariance build):
// Temme's series (x <= 2) or Steed's continued fraction (x > 2) for K_mu
// and K_(mu+1) as in xnu_bessel_k, stopped at FP32 accuracy, with x^mu (and
// e^-x) folded into the two starting values a_0 = x^mu K_mu(x),
// a_1 = x^(mu+1) K_(mu+1)(x).  The upward recurrence K_(v+1) = K_(v-1) +
// (2v/x) K_v multiplied by x^(v+1) reads a_(j+1) = 2 (mu + j) a_j +
// x^2 a_(j-1): it needs no 1/x and its terms stay finite as x -> 0, where
// K_nu alone overflows FP32.  Hardware exp2 / log2 / reciprocal (about
// 1 ulp each).
__device__ inline float xnu_bessel_k32(float x, const CovArgs32 &a,
                                       const BesselTab32 &t) {
  constexpr float eps = 6e-8f;
  constexpr float kLn2 = 0.69314718f, kLog2e = 1.44269504f;
  const float mu = a.mu;
  const float l2x = __builtin_amdgcn_logf(x);  // log2 x
  float k0, k1;
  if (x <= 2.0f) {
    const float lg = (1.0f - l2x) * kLn2;  // -ln(x / 2)
    const float sg = mu * lg;
    const float es = __builtin_amdgcn_exp2f(sg * kLog2e);
    const float ies = __builtin_amdgcn_rcpf(es);
    float fs;  // sinh(sg) / sg
    if (fabsf(sg) < 0.1f) {
      const float s2 = sg * sg;
      fs = 1.0f + s2 * (1.0f / 6.0f + s2 * (1.0f / 120.0f + s2 * (1.0f / 5040.0f)));
    } else {
      fs = 0.5f * (es - ies) * __builtin_amdgcn_rcpf(sg);
    }
    float f = a.fa * (a.g1 * (0.5f * (es + ies)) + a.g2 * fs * lg);
    float pk = 0.5f * es * a.gp;
    float qk = 0.5f * ies * a.gm;
    float ck = 1.0f;
    const float d = 0.25f * x * x;
    float s0 = f, s1 = pk;
    for (int k = 1; k < 100; k++) {
      float i2, i1, im, ip;
      const float fk = (float)k;
      if (k < kBesTab32) {
        i2 = t.ik2[k];
        i1 = t.ik[k];
        im = t.ikm[k];
        ip = t.ikp[k];
      } else {
        i2 = __builtin_amdgcn_rcpf(fk * fk - mu * mu);
        i1 = __builtin_amdgcn_rcpf(fk);
        im = __builtin_amdgcn_rcpf(fk - mu);
        ip = __builtin_amdgcn_rcpf(fk + mu);
      }
      f = (fk * f + pk + qk) * i2;
      ck *= d * i1;
      pk *= im;
      qk *= ip;
      const float t0 = ck * f;
      s0 += t0;
      s1 += ck * (pk - fk * f);
      if (fabsf(t0) < fabsf(s0) * eps) break;
    }
    const float xmu = __builtin_amdgcn_exp2f(mu * l2x);  // x^mu
    k0 = s0 * xmu;
    k1 = 2.0f * s1 * xmu;  // x^(mu+1) (2 / x) s1
  } else {
    float b = 2.0f * (1.0f + x);
    float dd = __builtin_amdgcn_rcpf(b);
    float h = dd, dh = dd;
    float q1 = 0.0f, q2 = 1.0f;
    const float a1 = 0.25f - mu * mu;
    float q = a1, c = a1;
    float aa = -a1;
    float s = 1.0f + q * dh;
    for (int i = 2; i < 1000; i++) {
      aa -= 2.0f * (float)(i - 1);
      float inv_i, inv_aa;
      if (i < kBesTab32) {
        inv_i = t.ik[i];
        inv_aa = t.iaa[i];
      } else {
        inv_i = __builtin_amdgcn_rcpf((float)i);
        inv_aa = __builtin_amdgcn_rcpf(aa);
      }
      c = -aa * c * inv_i;
      const float qn = (q1 - b * q2) * inv_aa;
      q1 = q2;
      q2 = qn;
      q += c * qn;
      b += 2.0f;
      dd = __builtin_amdgcn_rcpf(b + aa * dd);
      dh = (b * dd - 1.0f) * dh;
      h += dh;
      const float ds = q * dh;
      s += ds;
      if (fabsf(ds) < fabsf(s) * eps) break;
    }
    h = a1 * h;
    // x^mu K_mu = sqrt(pi / 2) x^(mu - 1/2) e^-x / s
    k0 = 1.25331414f *
         __builtin_amdgcn_exp2f((mu - 0.5f) * l2x - x * kLog2e) *
         __builtin_amdgcn_rcpf(s);
    k1 = k0 * (mu + x + 0.5f - h);
  }
  const float x2 = x * x;
  for (int j = 1; j <= a.nl; j++) {
    const float kn = fmaf(2.0f * (mu + (float)j), k1, x2 * k0);
    k0 = k1;
    k1 = kn;
  }
  return k0;
}

// s2 * rho in FP32 for any covModel (the matern through xnu_bessel_k32),
// from an FP32 squared distance.
__device__ inline float s2_corr32_rt(float d2, const CovArgs32 &a,
                                     const BesselTab32 &t) {
  switch (a.cov) {
    case 0: return s2_corr32<0>(d2, a);
    case 1: return s2_corr32<1>(d2, a);
    case 2: {
      const float x = __builtin_amdgcn_sqrtf(d2) * a.phi;
      if (!(x > 0.0f)) return a.s2;
      if (!(x < 80.0f)) return 0.0f;  // below 1e-30 of s2
      return a.mnorm * xnu_bessel_k32(x, a, t);
    }
    default: return s2_corr32<3>(d2, a);
  }
}

// FP32 Cholesky of a packed lower triangle (diagonal kept as 1/L_jj);
// false when a pivot is not positive.
template <int M>
__device__ __forceinline__ bool chol_factor32(float (&a)[M * (M + 1) / 2]) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < M; j++) {
    const float piv = a[tri(j, j)];
    ok = ok && (piv > 0.0f);
    const float r = __builtin_amdgcn_rsqf(piv);
    a[tri(j, j)] = r;
#pragma unroll
    for (int k = j + 1; k < M; k++) a[tri(k, j)] *= r;
#pragma unroll
    for (int k = j + 1; k < M; k++) {
      const float lkj = a[tri(k, j)];
#pragma unroll
      for (int l = j + 1; l <= k; l++)
        a[tri(k, l)] = fmaf(-lkj, a[tri(l, j)], a[tri(k, l)]);
    }
  }
  return ok;
}

// v <- C^-1 v with the factor of chol_factor32.
template <int M>
__device__ __forceinline__ void chol_apply32(const float (&a)[M * (M + 1) / 2],
                                             float (&v)[M]) {
#pragma unroll
  for (int k = 0; k < M; k++) {
    float t = v[k];
#pragma unroll
    for (int l = 0; l < k; l++) t = fmaf(-a[tri(k, l)], v[l], t);
    v[k] = t * a[tri(k, k)];
  }
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {
    const float t = v[k] * a[tri(k, k)];
    v[k] = t;
#pragma unroll
    for (int l = 0; l < k; l++) v[l] = fmaf(-a[tri(k, l)], t, v[l]);
  }
}

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of the two 32x32->64 multiplies with
// the published multipliers and Weyl key increments.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0,
                                              uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Standard normal keyed by (seed, site i, sample s): Philox4x32-10 of the
// counter (i, s, 0, 0) under key seed, two 53-bit uniforms in (0, 1), and
// the cosine branch of Box-Muller.  A draw depends only on its key and
// counter, so a kriging request gives the same draws however its sites are
// split into launches.
__host__ __device__ inline double philox_normal(uint64_t seed, uint32_t i,
                                                uint32_t s) {
  uint32_t c[4] = {i, s, 0u, 0u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((uint64_t)c[0] << 32) | c[1];
  const uint64_t b = ((uint64_t)c[2] << 32) | c[3];
  const double u1 = ((double)(a >> 11) + 0.5) * 0x1p-53;
  const double u2 = ((double)(b >> 11) + 0.5) * 0x1p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X; give each
// XCD a contiguous range of logical blocks so that neighbor gathers of one
// XCD stay in its own L2.  grid = 8 * per; returns -1 for padding blocks.
__device__ __forceinline__ int xcd_block(int nlogical) {
  int per = (nlogical + 7) >> 3;
  int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  return b < nlogical ? b : -1;
}

// The sites of wave w of logical block b (from xcd_block) in a launch whose
// nlogical blocks of nw waves each stay resident while they loop over n
// sites.  XCD x owns the contiguous run of its logical blocks' shares, and
// the waves of those blocks take the run's sites round robin: at any moment
// they work on neighboring sites, whose neighbor gathers then share the
// XCD's L2.  (A contiguous share per block would spread the XCD's waves
// over its whole run.)  The wave takes first, first + stride, ... < end.
struct SiteRun {
  int first, stride, end;
};
__device__ __forceinline__ SiteRun xcd_site_run(int nlogical, int b, int n,
                                                int nw, int w) {
  const int per = (nlogical + 7) >> 3;
  const int x = b / per, j = b - x * per;
  const int nb = min(per, nlogical - x * per);  // logical blocks on XCD x
  const long long chunk = ((long long)n + nlogical - 1) / nlogical;
  const long long s0 = (long long)x * per * chunk;
  const long long s1 = min((long long)n, s0 + (long long)nb * chunk);
  SiteRun r;
  r.first = (int)min(s0 + (long long)j * nw + w, s1);
  r.stride = nb * nw;
  r.end = (int)s1;
  return r;
}

}  // namespace nngp
