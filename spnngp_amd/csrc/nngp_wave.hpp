// nngp_wave.hpp — wavefront-level pieces shared by the one-system-per-
// wavefront kernels (updateBF, kriging, conjugate and k-fold prediction):
// wave synchronisation, lane broadcasts and sums, and the Cholesky solve of
// one neighbor system with one row per lane.
#pragma once
#include <hip/hip_runtime.h>

#include "nngp_device.hpp"

namespace nngp {

// Orders the wave's LDS writes before its later LDS reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The lane index, opaque to the optimizer.  The unrolled factor and solves
// below compare it with every row number (t > k, t == k, l <= t); where the
// caller loops over sites, loop-invariant code motion would otherwise hoist
// all of those compares out of the loop as 64-bit lane masks, more than the
// SGPR file holds, and spill them to VGPR lanes (a readlane per use).
__device__ __forceinline__ int lane_opaque(int t) {
  asm volatile("" : "+v"(t));
  return t;
}

// Lane `lane`'s v, in every lane (lane must be wave-uniform).
__device__ __forceinline__ double readlane_d(double v, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Sum over the 64 lanes by a butterfly, identical in every lane.
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Cholesky of the cnt x cnt lower triangle in LDS (row-major, ld; row t
// read by lane t) and the two triangular solves, for cnt <= MM <= 64.
// Lane t keeps row t of the factor in registers; the entries of column j
// reach the other rows by readlane, so the right-looking update is register
// arithmetic (no LDS round trips in the O(m^3) part).  The factor goes back
// to LDS for the back substitution, which needs its columns.  ct is lane
// t's right-hand side; on return x = (C^-1 c)_t for t < cnt (lanes t >= cnt
// get ct back) and vv = c^T C^-1 c in every lane.  False (uniformly) when a
// pivot is not positive.  Every lane of the wave must call it.
template <int MM>
__device__ inline bool wave_chol_rows(double *Am, int ld, int cnt, int t,
                                      double ct, double &x, double &vv) {
  t = lane_opaque(t);
  // Rows t >= cnt stay zero, so their L_tj is zero and the updates below
  // need no bound on l.  Entries above lane t's diagonal collect values
  // that no step reads: step j broadcasts the L_lj of lanes l > j only.
  double row[MM];
#pragma unroll
  for (int l = 0; l < MM; l++)
    row[l] = (t < cnt && l <= t) ? Am[t * ld + l] : 0.0;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < MM; j++) {
    // a guard, not a break: a loop exit would stop the full unrolling that
    // keeps row[] in registers
    if (j < cnt) {
      const double piv = readlane_d(row[j], j);
      ok = ok && (piv > 0.0);
      const double r = nn_rsqrt(piv > 0.0 ? piv : 1.0);
      const double ltj = row[j] * r;  // L_tj for t > j
      row[j] = t == j ? r : ltj;      // the diagonal keeps 1 / L_jj
#pragma unroll
      for (int l = j + 1; l < MM; l++)
        row[l] = fma(-ltj, readlane_d(ltj, l), row[l]);
    }
  }
  if (!ok) return false;
  // forward solve: lane t finishes x_t at step t
  double xt = ct;
#pragma unroll
  for (int k = 0; k < MM; k++) {
    if (k < cnt) {
      const double xk = readlane_d(xt * row[k], k);
      xt = t == k ? xk : (t > k ? fma(-row[k], xk, xt) : xt);
    }
  }
  vv = wave_sum(t < cnt ? xt * xt : 0.0);
#pragma unroll
  for (int l = 0; l < MM; l++)
    if (t < cnt && l <= t) Am[t * ld + l] = row[l];
  wave_sync();
  // back solve B = L^-T x: column k of the factor is row k's entries
  for (int k = cnt - 1; k >= 0; k--) {
    const double bk = readlane_d(xt, k) * Am[k * ld + k];
    if (t == k) xt = bk;
    if (t < k) xt = fma(-Am[k * ld + t], bk, xt);
  }
  x = xt;
  return true;
}

// FP32 Cholesky of a cnt x cnt matrix held one row per lane: lane t passes
// row t's entries l <= t in row[] (zeros elsewhere, all zeros for t >= cnt)
// and gets back row t of the factor, the diagonal as 1 / L_tt (entries
// above the diagonal are left unused, as in wave_chol_rows).  False
// (uniformly) when a pivot is not positive.
template <int MM>
__device__ inline bool wave_factor32(float (&row)[MM], int cnt, int t) {
  t = lane_opaque(t);
  bool ok = true;
#pragma unroll
  for (int j = 0; j < MM; j++) {
    if (j < cnt) {  // a guard, not a break (see wave_chol_rows)
      const float piv = readlane_f(row[j], j);
      ok = ok && (piv > 0.0f);
      const float r = __builtin_amdgcn_rsqf(piv > 0.0f ? piv : 1.0f);
      const float ltj = row[j] * r;
      row[j] = t == j ? r : ltj;
#pragma unroll
      for (int l = j + 1; l < MM; l++)
        row[l] = fmaf(-ltj, readlane_f(ltj, l), row[l]);
    }
  }
  return ok;
}

// y = C^-1 b in FP32 with the factor of wave_factor32: the forward
// substitution from the rows in registers, the backward one from the same
// factor stored row-major in LDS (Lm, ld; entries l <= k of row k).  Lane t
// holds b_t and gets y_t (lanes t >= cnt get b_t back).
template <int MM>
__device__ inline float wave_solve32(const float (&row)[MM], const float *Lm,
                                     int ld, int cnt, int t, float b) {
  t = lane_opaque(t);
  float y = b;
#pragma unroll
  for (int k = 0; k < MM; k++) {
    if (k < cnt) {
      const float yk = readlane_f(y * row[k], k);
      y = t == k ? yk : (t > k ? fmaf(-row[k], yk, y) : y);
    }
  }
  for (int k = cnt - 1; k >= 0; k--) {
    const float xk = readlane_f(y, k) * Lm[k * ld + k];
    if (t == k) y = xk;
    if (t < k) y = fmaf(-Lm[k * ld + t], xk, y);
  }
  return y;
}

// Row count of the register factor for a system of m <= 64 rows.
__host__ __device__ constexpr int wave_rows_for(int m) {
  return m <= 16 ? 16 : (m <= 32 ? 32 : 64);
}

}  // namespace nngp
