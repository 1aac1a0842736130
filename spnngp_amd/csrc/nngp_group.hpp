// nngp_group.hpp — one neighbor system per 16-lane row of the wavefront
// (four sites per wave), for the general-nu Matern and m > 16, where one
// site per lane does not fit in registers.
//
// Lane u of a row holds rows u and u + 16 of its site's MM x MM system
// (MM = 16 or 32), so all 64 lanes carry matrix rows; a wavefront-wide
// system (one site per wave) leaves half of them idle at m = 32.  Within a
// row, DPP row_newbcast moves a double from lane l to the other 15 lanes
// in one full-rate instruction (v_mov_b64_dpp), which serves the row's
// site and, at once, the other three rows' sites: the factor's broadcast
// of L_lj costs one instruction per four systems instead of two
// v_readlane_b32 per system.  The row's sums (vv, the back substitution's
// dot products, the residuals) are butterflies of four DPP steps.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "nngp_wave.hpp"

namespace nngp {

// f(std::integral_constant<int, J>) for J = B .. E-1, each J a constant
// expression in f's body (the DPP lane selects need one).
template <int B, int E, class Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Lane L of each 16-lane row's v, in every lane of that row.  Every lane
// has a valid source, so no old value is prepared (bound_ctrl).
template <int L>
__device__ __forceinline__ double row_bcast(double v) {
  static_assert(L >= 0 && L < 16, "lane within a row");
  return __builtin_amdgcn_update_dpp(0.0, v, 0x150 + L, 0xF, 0xF, true);
}

// One 32-bit DPP move of each half of a double, for row-local patterns in
// which every lane has a valid source.
template <int CTRL>
__device__ __forceinline__ double row_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF,
                                             0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF,
                                             0xF, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of each row, in every lane of the row: pairs, quads
// (quad_perm), eights (row_half_mirror), sixteen (row_mirror).  The order is
// fixed, so the sum is the same in every lane and every run.
__device__ __forceinline__ double row_sum16(double x) {
  x += row_dpp<0xB1>(x);   // quad_perm [1,0,3,2]
  x += row_dpp<0x4E>(x);   // quad_perm [2,3,0,1]
  x += row_dpp<0x141>(x);  // row_half_mirror
  x += row_dpp<0x140>(x);  // row_mirror
  return x;
}

// The system of one site, two rows a lane (the second unused at MM = 16).
// ra[l] is row u, rb[l] row u + 16; entries above a row's diagonal are
// scratch.  After grp_factor the diagonal holds 1 / L_kk.
template <int MM>
struct GrpRows {
  static_assert(MM == 16 || MM == 32, "16 or 32 rows");
  static constexpr int R = MM / 16;
  double ra[16];
  double rb[R == 2 ? 32 : 1];
};

// Right-looking Cholesky of the row's system (rows k >= the site's count
// are identity rows, so no step needs a bound).  u is the lane within the
// row.  True when every pivot is positive.
template <int MM>
__device__ __forceinline__ bool grp_factor(GrpRows<MM> &S, int u) {
  constexpr int R = GrpRows<MM>::R;
  bool ok = true;
  static_for<0, MM>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    double piv;
    if constexpr (j < 16) piv = row_bcast<j>(S.ra[j]);
    else piv = row_bcast<j - 16>(S.rb[j]);
    ok = ok && (piv > 0.0);
    const double r = nn_rsqrt(piv > 0.0 ? piv : 1.0);
    double la = 0.0, lb = 0.0;
    if constexpr (j < 16) {
      la = S.ra[j] * r;               // L_uj for u > j
      S.ra[j] = u == j ? r : la;      // the diagonal keeps 1 / L_jj
    }
    if constexpr (R == 2) {
      lb = S.rb[j] * r;               // L_(u+16)j
      S.rb[j] = u + 16 == j ? r : lb;
    }
    if constexpr (j < 15) {
      static_for<j + 1, 16>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const double L = row_bcast<l>(la);  // L_lj from lane l
        S.ra[l] = fma(-la, L, S.ra[l]);
        if constexpr (R == 2) S.rb[l] = fma(-lb, L, S.rb[l]);
      });
    }
    if constexpr (R == 2) {
      static_for<(j + 1 > 16 ? j + 1 : 16), 32>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const double L = row_bcast<l - 16>(lb);  // L_lj from lane l - 16
        S.rb[l] = fma(-lb, L, S.rb[l]);
      });
    }
  });
  return ok;
}

// y = L^-1 (ca, cb) (the lane's rows u and u + 16 of the right-hand side);
// returns the lane's (y_u, y_u+16).
template <int MM>
__device__ __forceinline__ void grp_forward(const GrpRows<MM> &S, int u,
                                            double ca, double cb, double &ya,
                                            double &yb) {
  constexpr int R = GrpRows<MM>::R;
  double aa = ca, ab = cb;
  ya = 0.0;
  yb = 0.0;
  static_for<0, MM>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if constexpr (k < 16) {
      const double yk = row_bcast<k>(aa * S.ra[k]);  // final in lane k
      ya = u == k ? yk : ya;
      aa = fma(-S.ra[k], yk, aa);
      if constexpr (R == 2) ab = fma(-S.rb[k], yk, ab);
    } else {
      const double yk = row_bcast<k - 16>(ab * S.rb[k]);
      yb = u + 16 == k ? yk : yb;
      ab = fma(-S.rb[k], yk, ab);
    }
  });
}

// B = L^-T y: going down from the last row, the dot product of column k
// below the diagonal with the B already final is a row sum.  Returns the
// lane's (B_u, B_u+16).
template <int MM>
__device__ __forceinline__ void grp_backward(const GrpRows<MM> &S, int u,
                                             double ya, double yb, double &ba,
                                             double &bb) {
  constexpr int R = GrpRows<MM>::R;
  ba = 0.0;
  bb = 0.0;
  static_for<0, MM>([&](auto qc) {
    constexpr int k = MM - 1 - decltype(qc)::value;
    double s = 0.0;
    if constexpr (k < 16) s = u > k ? S.ra[k] * ba : 0.0;
    if constexpr (R == 2) s += u + 16 > k ? S.rb[k] * bb : 0.0;
    s = row_sum16(s);
    if constexpr (k < 16) ba = u == k ? (ya - s) * S.ra[k] : ba;
    else bb = u + 16 == k ? (yb - s) * S.rb[k] : bb;
  });
}

}  // namespace nngp
