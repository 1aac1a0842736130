// nngp_rrng.hip — R's default random number generator on the host, so that a
// chain run through the library without R follows the stream R's
// set.seed(seed) would give: RNGkind "Mersenne-Twister" with normal.kind
// "Inversion" (R's defaults since 1.7.0).
//
//  * Seeding (set.seed): the integer seed passes 50 times through the
//    linear congruential map s <- 69069 s + 1 (mod 2^32), then the next
//    625 values of the same map fill the generator's seed vector; its first
//    entry (the position in the state) is then set to 624, so the first
//    draw regenerates the whole state.
//  * Uniforms: MT19937 (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998) 32-bit
//    outputs times 2^-32, kept inside (0, 1).
//  * Normals by inversion: u = floor(2^27 U1) + U2, z = Phi^-1(u / 2^27)
//    with Phi^-1 from Wichura's algorithm AS 241 (PPND16, Applied
//    Statistics 37, 1988), accurate to about 1e-16.
// The generator lives only on the host; nngp_rng carries it into the
// samplers like any caller-supplied source.
#include <cmath>
#include <cstdint>
#include <new>

#include "../../include/nngp.h"

namespace {

struct RState {
  uint32_t mt[624];
  int mti = 625;
};

constexpr int kN = 624, kM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpper = 0x80000000u,
                   kLower = 0x7fffffffu;

uint32_t mt_next(RState *s) {
  static const uint32_t mag01[2] = {0u, kMatrixA};
  if (s->mti >= kN) {
    if (s->mti == kN + 1) {  // never seeded: MT's default initial state
      uint32_t seed = 4357u;
      for (int i = 0; i < kN; i++) {
        s->mt[i] = seed & 0xffff0000u;
        seed = 69069u * seed + 1u;
        s->mt[i] |= (seed & 0xffff0000u) >> 16;
        seed = 69069u * seed + 1u;
      }
    }
    int k = 0;
    for (; k < kN - kM; k++) {
      const uint32_t y = (s->mt[k] & kUpper) | (s->mt[k + 1] & kLower);
      s->mt[k] = s->mt[k + kM] ^ (y >> 1) ^ mag01[y & 1u];
    }
    for (; k < kN - 1; k++) {
      const uint32_t y = (s->mt[k] & kUpper) | (s->mt[k + 1] & kLower);
      s->mt[k] = s->mt[k + (kM - kN)] ^ (y >> 1) ^ mag01[y & 1u];
    }
    const uint32_t y = (s->mt[kN - 1] & kUpper) | (s->mt[0] & kLower);
    s->mt[kN - 1] = s->mt[kM - 1] ^ (y >> 1) ^ mag01[y & 1u];
    s->mti = 0;
  }
  uint32_t y = s->mt[s->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double r_unif(void *st) {
  const double i2_32m1 = 2.328306437080797e-10;  // 1 / (2^32 - 1)
  const double x = (double)mt_next(static_cast<RState *>(st)) *
                   2.3283064365386963e-10;  // 2^-32
  if (x <= 0.0) return 0.5 * i2_32m1;
  if (1.0 - x <= 0.0) return 1.0 - 0.5 * i2_32m1;
  return x;
}

// Phi^-1(p) for 0 < p < 1: AS 241 PPND16 (Wichura 1988), the central
// rational approximation for |p - 1/2| <= 0.425, the two tail ones beyond.
double ppnd16(double p) {
  const double q = p - 0.5;
  if (std::fabs(q) <= 0.425) {
    const double r = 0.180625 - q * q;
    return q *
           (((((((r * 2509.0809287301226727 + 33430.575583588128105) * r +
                 67265.770927008700853) * r + 45921.953931549871457) * r +
               13731.693765509461125) * r + 1971.5909503065514427) * r +
             133.14166789178437745) * r + 3.387132872796366608) /
           (((((((r * 5226.495278852854561 + 28729.085735721942674) * r +
                 39307.89580009271061) * r + 21213.794301586595867) * r +
               5394.1960214247511077) * r + 687.1870074920579083) * r +
             42.313330701600911252) * r + 1.0);
  }
  double r = q < 0.0 ? p : 1.0 - p;
  r = std::sqrt(-std::log(r));
  double val;
  if (r <= 5.0) {
    r -= 1.6;
    val = (((((((r * 7.7454501427834140764e-4 + 0.0227238449892691845833) * r +
                0.24178072517745061177) * r + 1.27045825245236838258) * r +
              3.64784832476320460504) * r + 5.7694972214606914055) * r +
            4.6303378461565452959) * r + 1.42343711074968357734) /
          (((((((r * 1.05075007164441684324e-9 + 5.475938084995344946e-4) *
                    r + 0.0151986665636164571966) * r +
                0.14810397642748007459) * r + 0.68976733498510000455) * r +
              1.6763848301838038494) * r + 2.05319162663775882187) * r + 1.0);
  } else {
    r -= 5.0;
    val = (((((((r * 2.01033439929228813265e-7 + 2.71155556874348757815e-5) *
                    r + 0.0012426609473880784386) * r +
                0.026532189526576123093) * r + 0.29656057182850489123) * r +
              1.7848265399172913358) * r + 5.4637849111641143699) * r +
            6.6579046435011037772) /
          (((((((r * 2.04426310338993978564e-15 + 1.4215117583164458887e-7) *
                    r + 1.8463183175100546818e-5) * r +
                7.868691311456132591e-4) * r + 0.0148753612908506148525) * r +
              0.13692988092273580531) * r + 0.59983220655588793769) * r + 1.0);
  }
  return q < 0.0 ? -val : val;
}

double r_norm(void *st) {
  const double big = 134217728.0;  // 2^27
  double u = r_unif(st);
  u = (double)(int)(big * u) + r_unif(st);
  return ppnd16(u / big);
}

}  // namespace

extern "C" int nngp_rng_r_create(unsigned int seed, nngp_rng *out) {
  if (!out) return NNGP_ERR_ARG;
  RState *s = new (std::nothrow) RState;
  if (!s) return NNGP_ERR_NOMEM;
  uint32_t x = seed;
  for (int j = 0; j < 50; j++) x = 69069u * x + 1u;  // initial scrambling
  // 625 more values: the first lands in the position slot, which is then
  // set to 624 (so the first draw regenerates the state); the rest are
  // the 624 state words
  x = 69069u * x + 1u;
  for (int j = 0; j < kN; j++) {
    x = 69069u * x + 1u;
    s->mt[j] = x;
  }
  s->mti = kN;
  out->state = s;
  out->norm_rand = r_norm;
  out->unif_rand = r_unif;
  return NNGP_OK;
}

extern "C" void nngp_rng_r_destroy(nngp_rng *rng) {
  if (!rng || !rng->state) return;
  delete static_cast<RState *>(rng->state);
  rng->state = nullptr;
}
