// nngp_latent.hip — the latent NNGP model: the neighbors-of-neighbors index
// (mkUIndx, src/nn.cpp:25-56) and the Gibbs / Metropolis sampler sNNGP
// (src/sNNGP.cpp:90-482).
//
// mkUIndx on the device.  For each site c the sites r that hold c as a
// neighbor are the transpose of the neighbor lists.  Every (r, slot) entry
// is enumerated in row-major order; a stable radix sort on the neighbor id
// (hipCUB) groups the entries by c while keeping ascending r inside each
// group, which is the order of the reference's scans 0 and 2 (search type 1
// lists the group backwards).  The slot of c in r's list (uiIndx) falls out
// of the entry's position.
//
// The sampler keeps everything resident: coordinates, neighbor and
// transpose indices, y, X, the two B/F slots and w (stored as the one data
// column of a Model, so that each updateBF pass also yields
// sum_i (w_i - B_i w_N(i))^2 / F_i and sum_i log F_i).  The w sweep updates
// every site from its full conditional.  Its order is a schedule of phases:
//  * NNGP_SWEEP_SITE_ORDER: the reference's order 0..n-1.  Site i reads the
//    current w of its neighbors, of the sites holding it as a neighbor and
//    of their other neighbors (its neighbors in the moral graph); sites
//    grouped by dependency level (each level above every adjacent
//    predecessor) can be updated together, and a level-by-level sweep
//    reads exactly the values the sequential sweep reads.
//  * NNGP_SWEEP_COLOURED: a greedy colouring of the moral graph; one phase
//    per colour.  This is a Gibbs scan in another site order (colour by
//    colour), with the same stationary distribution.
// Phases with many sites get a grid of their own; runs of small phases go
// to one block that walks them with a barrier between phases.  Each site
// reads its children's residuals w_j - B_j . w_N(j), kept current through
// the sweep, instead of re-forming them from the children's neighbor lists
// (sweep_site).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nngp_host.hpp"

namespace nngp {

// ---------------------------------------------------------------------------
// mkUIndx
// ---------------------------------------------------------------------------
constexpr int kUThreads = 256;

// Row r's entries: key = neighbor id, value = row-major position; owner
// maps a position back to its row; ucnt counts each neighbor id.
__global__ __launch_bounds__(kUThreads) void k_u_enumerate(
    const int *nnIndx, const int *lu, const int *rowstart, int n, int *keys,
    int *vals, int *owner, int *ucnt) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= n) return;
  const int off = lu[r], cnt = lu[n + r], base = rowstart[r];
  for (int s = 0; s < cnt; s++) {
    const int c = nnIndx[off + s];
    keys[base + s] = c;
    vals[base + s] = base + s;
    owner[base + s] = r;
    atomicAdd(&ucnt[c], 1);
  }
}

// Sorted entry t of group c: its row and slot, at t (ascending) or mirrored
// inside the group (descending).
__global__ __launch_bounds__(kUThreads) void k_u_place(
    const int *skeys, const int *svals, const int *owner, const int *rowstart,
    const int *uoff, const int *ucnt, int total, int descending, int *uIndx,
    int *uiIndx) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= total) return;
  const int c = skeys[t], pos = svals[t];
  const int r = owner[pos];
  const int dst = descending ? 2 * uoff[c] + ucnt[c] - 1 - t : t;
  uIndx[dst] = r;
  uiIndx[dst] = pos - rowstart[r];
}

// ---------------------------------------------------------------------------
// w sweep
// ---------------------------------------------------------------------------
constexpr int kSweepBlock = 1024;  // the block that walks small phases
constexpr int kSweepGrid = 256;    // threads per block of a phase's grid
constexpr int kGridPhase = 4096;   // phases at least this large get a grid

struct SweepArgs {
  const double *B, *F, *y, *X, *z;
  double *w;
  const int *nnIndx, *uIndx, *uiIndx, *phase;
  // per schedule position k: {site i, uIndxLU[i], uIndxLU[n+i], 0}; per
  // transpose entry t: {nnIndxLU[j], nnIndxLU[n+j]} of the site j = uIndx[t]
  // (the per-site lookups of the sweep, laid out in the order it reads them)
  const int4 *meta;
  const int2 *uoc;
  // per transpose entry t: B of the site uIndx[t] at its neighbor slot that
  // names the entry's site, gathered from B once per B (k_gather_bt)
  const double *Bt;
  const int *lu;  // nnIndxLU
  // per site {F_i, r_i}: F of the sweep's B, F slot and the residual
  // w_i - B_i . w_N(i) at the current w, side by side so that a child's two
  // values come in one 16 B read
  double2 *FR;
  double beta[kMaxP];
  double tau2;
  unsigned long long seed;  // Philox key when z == nullptr
  int iter;                 // Philox counter word of this sweep
  int n, p, ws;             // w[i * ws]
  int ph0, ph1;             // phases [ph0, ph1)
  int nph;                  // phases in the schedule (phase[] has nph + 1)
  // k_resid_init: F of the slot times this, the slot's F at the current
  // sigma.sq (F is proportional to sigma.sq without a nugget)
  double fscale = 1.0;
};

// sum_k B[o+k] w[nnIndx[o+k]] over k < cnt, skipping the entry equal to
// `skip`, in k order.  The index and coefficient loads of a chunk of
// kChunk entries, then its w loads, are all issued before the first use, so
// a chunk costs two memory latencies instead of one per entry.
constexpr int kChunk = 16;

__device__ __forceinline__ double gather_dot(const SweepArgs &A, int o,
                                             int cnt, int skip) {
#pragma clang fp contract(off)
  double acc = 0.0;
  for (int k0 = 0; k0 < cnt; k0 += kChunk) {
    int kk[kChunk];
    double bk[kChunk], wk[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; k++) {
      kk[k] = k0 + k < cnt ? A.nnIndx[o + k0 + k] : skip;
      bk[k] = k0 + k < cnt ? A.B[o + k0 + k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kChunk; k++)
      wk[k] = kk[k] != skip ? A.w[(size_t)kk[k] * A.ws] : 0.0;
#pragma unroll
    for (int k = 0; k < kChunk; k++)
      if (kk[k] != skip) acc += bk[k] * wk[k];
  }
  return acc;
}

// w_i from its full conditional (src/sNNGP.cpp:260-287): precision
// 1/tau.sq + 1/F_i + sum_j B_ji^2/F_j over the sites j holding i as a
// neighbor (its children); mean from the data, i's own neighbor sum and
// each child's residual without i's term.
//
// The sweep keeps r_j = w_j - B_j . w_N(j) of every site at the current w
// (k_resid_init sets them before a sweep).  Child j's residual without i is
// then r_j + B_ji w_i and i's own neighbor sum is w_i - r_i, so a site
// reads one value per child instead of the child's whole neighbor list.
// After the draw, d = w_i(new) - w_i(old) moves r_i by +d and each child's
// r_j by -B_ji d.  No two sites of a phase are neighbors in the moral graph,
// so within a phase no two sites share a child and these updates never
// collide.  (The reference forms each child's sum afresh; the two agree to
// rounding.)
//
// A group of kGroup lanes takes one site; its children are taken kChild at
// a time, lane l holding children l, l + kGroup, l + 2 kGroup of the chunk,
// so every load of a chunk is in flight at once.  Every lane adds the
// children's terms in child order (shuffles) and forms the same draw.  All
// lanes of a group run the same trip counts.
constexpr int kGroup = 8;
constexpr int kPer = 3;
constexpr int kChild = kGroup * kPer;

__device__ void sweep_site(const SweepArgs &A, int4 mt, int lane) {
  const int n = A.n, ws = A.ws;
  double *w = A.w;
  const int i = mt.x, uo = mt.y, uc = mt.z;
  const double wi = w[(size_t)i * ws];
  const double2 fri = A.FR[i];
  const double fi = fri.x, ri = fri.y;
  const double yi = A.y[i];
  double xb = 0.0;
  for (int q = 0; q < A.p; q++) xb += A.X[(size_t)q * n + i] * A.beta[q];
  double a = 0.0, v = 0.0;
  // the first chunk's children (index, B_ji, r_j), kept for the residual
  // updates; no other site of the phase touches r_j, so the value read here
  // is still current when it is updated
  int jk[kPer];
  double bk[kPer], rk[kPer];
  for (int j0 = 0; j0 < uc; j0 += kChild) {
    double ta[kPer], tv[kPer];
#pragma unroll
    for (int c = 0; c < kPer; c++) {
      ta[c] = 0.0;
      tv[c] = 0.0;
      const int j = j0 + c * kGroup + lane;
      if (j < uc) {
        const int t = uo + j;
        const int jj = A.uIndx[t];
        const double bij = A.Bt[t];
        const double2 frj = A.FR[jj];
        const double fj = frj.x, rj = frj.y;
        ta[c] = bij * (rj + bij * wi) / fj;
        tv[c] = bij * bij / fj;
        if (j0 == 0) {
          jk[c] = jj;
          bk[c] = bij;
          rk[c] = rj;
        }
      }
    }
    const int nt = min(kChild, uc - j0);
#pragma unroll
    for (int t = 0; t < kChild; t++) {
      if (t < nt) {
        a += __shfl(ta[t / kGroup], t % kGroup, kGroup);
        v += __shfl(tv[t / kGroup], t % kGroup, kGroup);
      }
    }
  }
  const double e = wi - ri;
  const double mu = (yi - xb) / A.tau2 + e / fi + a;
  const double var = 1.0 / (1.0 / A.tau2 + 1.0 / fi + v);
  const double zi =
      A.z ? A.z[i] : philox_normal(A.seed, (uint32_t)i, (uint32_t)A.iter);
  const double wn = mu * var + sqrt(var) * zi;
  const double d = wn - wi;
  if (lane == 0) {
    w[(size_t)i * ws] = wn;
    A.FR[i].y = ri + d;
  }
  for (int j0 = 0; j0 < uc; j0 += kChild) {
#pragma unroll
    for (int c = 0; c < kPer; c++) {
      const int j = j0 + c * kGroup + lane;
      if (j >= uc) continue;
      int jj;
      double bij, rj;
      if (j0 == 0) {
        jj = jk[c];
        bij = bk[c];
        rj = rk[c];
      } else {
        const int t = uo + j;
        jj = A.uIndx[t];
        bij = A.Bt[t];
        rj = A.FR[jj].y;
      }
      A.FR[jj].y = rj - bij * d;
    }
  }
}

// Bt[t] = B[bpos[t]]: the sweep's B entries in the order it reads them, so
// that a child's coefficient comes with its index instead of after two
// index loads.  Once per B (the sampler's B changes when a phi proposal is
// accepted).
__global__ __launch_bounds__(kSweepGrid) void k_gather_bt(double *Bt,
                                                         const double *B,
                                                         const int *bpos,
                                                         long long nt) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nt) Bt[t] = B[bpos[t]];
}

// The sweep's counter-based normals for every site at once: the site-order
// sweep then reads z_i with its other loads instead of computing the
// Philox draw (ten rounds, a log, a sqrt and a cos) on each level's
// critical path.  The same values as philox_normal in the sweep.
__global__ __launch_bounds__(kSweepGrid) void k_philox_fill(
    double *z, int n, unsigned long long seed, int iter) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) z[i] = philox_normal(seed, (uint32_t)i, (uint32_t)iter);
}

// {F_i, r_i = w_i - B_i . w_N(i)} for every site, before a sweep
__global__ __launch_bounds__(kSweepGrid) void k_resid_init(SweepArgs A) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= A.n) return;
  A.FR[i] = make_double2(A.F[i] * A.fscale,
                         A.w[(size_t)i * A.ws] -
                             gather_dot(A, A.lu[i], A.lu[A.n + i], -1));
}

// One block walks phases [ph0, ph1) with a barrier between them.  The
// boundaries of the phase after next and the first site of the next phase
// are read one phase ahead, so a phase starts with its sites known.
__global__ __launch_bounds__(kSweepBlock) void k_sweep_block(SweepArgs A) {
  const int g = (int)threadIdx.x / kGroup, lane = (int)threadIdx.x % kGroup;
  const int ng = (int)blockDim.x / kGroup;
  int a = A.phase[A.ph0], b = A.phase[A.ph0 + 1];
  int4 first = a + g < b ? A.meta[a + g] : make_int4(0, 0, 0, 0);
  for (int L = A.ph0; L < A.ph1; L++) {
    const int c = L + 2 <= A.nph ? A.phase[L + 2] : b;
    const int4 cur = first;
    if (b + g < A.n) first = A.meta[b + g];  // checked against c at use
    for (int k = a + g; k < b; k += ng)
      sweep_site(A, k == a + g ? cur : A.meta[k], lane);
    __syncthreads();
    a = b;
    b = c;
    if (a + g >= b) first = make_int4(0, 0, 0, 0);
  }
}

__global__ __launch_bounds__(kSweepGrid) void k_sweep_grid(SweepArgs A) {
  const int a = A.phase[A.ph0], b = A.phase[A.ph0 + 1];
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = a + t / kGroup;
  if (k < b) sweep_site(A, A.meta[k], t % kGroup);
}

// ---------------------------------------------------------------------------
// reductions: per-block partial sums, then one block adds the partials in
// block order (deterministic)
// ---------------------------------------------------------------------------
constexpr int kRedBlock = 256;
constexpr int kRedMax = 2048;  // partial blocks
constexpr int kLatStats = kMaxP + 1;

struct LatSumArgs {
  const double *y, *X, *w;
  const double2 *FR;  // {F_i, w_i - B_i . w_N(i)}, current after a sweep
  double beta[kMaxP];
  double tau2;
  int n, p, ws, mode;  // mode 0: X^T (y - w) / tau.sq and the w quadratic
                       // form sum_i r_i^2 / F_i; mode 1: ||y - w - X beta||^2
  double *part;        // [kRedMax][kLatStats]
};

__device__ __forceinline__ double block_sum(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  const int t = (int)threadIdx.x;
  if ((t & 63) == 0) red[t >> 6] = x;
  __syncthreads();
  double s = 0.0;
  if (t == 0)
    for (int k = 0; k < (int)blockDim.x / 64; k++) s += red[k];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(kRedBlock) void k_lat_partial(LatSumArgs A) {
  __shared__ double red[kRedBlock / 64];
  double acc[kLatStats];
  for (int k = 0; k < kLatStats; k++) acc[k] = 0.0;
  const int n = A.n;
  for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n;
       i += (int)(gridDim.x * blockDim.x)) {
    const double wi = A.w[(size_t)i * A.ws];
    if (A.mode == 0) {
      const double r = (A.y[i] - wi) / A.tau2;
      for (int k = 0; k < A.p; k++) acc[k] += A.X[(size_t)k * n + i] * r;
      const double2 fr = A.FR[i];
      acc[kMaxP] += fr.y * fr.y / fr.x;
    } else {
      double xb = 0.0;
      for (int k = 0; k < A.p; k++) xb += A.X[(size_t)k * n + i] * A.beta[k];
      const double r = A.y[i] - wi - xb;
      acc[0] += r * r;
    }
  }
  for (int k = 0; k < kLatStats; k++) {
    const double s = block_sum(acc[k], red);
    if (threadIdx.x == 0) A.part[(size_t)blockIdx.x * kLatStats + k] = s;
  }
}

__global__ __launch_bounds__(kRedBlock) void k_lat_final(const double *part,
                                                         int nblk,
                                                         double *out) {
  __shared__ double red[kRedBlock / 64];
  for (int k = 0; k < kLatStats; k++) {
    double s = 0.0;
    for (int b = (int)threadIdx.x; b < nblk; b += (int)blockDim.x)
      s += part[(size_t)b * kLatStats + k];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[k] = s;
  }
}

__global__ void k_gather_w(const double *w, int ws, int n, double *out) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) out[i] = w[(size_t)i * ws];
}

// ---------------------------------------------------------------------------
// host: schedules
// ---------------------------------------------------------------------------

// The transpose must name, for every child j of site i (j > i), a slot of
// j's list that holds i; anything else would make the sweep read outside
// B.  Also fixes that every neighbor precedes its site.
static int check_transpose(int n, const int *nnIndx, const int *lu,
                           const int *uIndx, const int *ulu,
                           const int *uiIndx, long long nIndx) {
  if (!uIndx || !ulu || !uiIndx)
    return set_err(NNGP_ERR_ARG, "uIndx/uIndxLU/uiIndx is NULL");
  long long tot = 0, ntot = 0;
  for (int i = 0; i < n; i++) {
    const long long o = ulu[i], c = ulu[(size_t)n + i];
    if (o < 0 || c < 0 || o + c > nIndx)
      return set_err(NNGP_ERR_ARG, "malformed uIndxLU at site %d", i);
    tot += c;
    ntot += lu[(size_t)n + i];
    for (long long j = 0; j < c; j++) {
      const int jj = uIndx[o + j], s = uiIndx[o + j];
      if (jj <= i || jj >= n || s < 0 || s >= lu[(size_t)n + jj] ||
          nnIndx[lu[jj] + s] != i)
        return set_err(NNGP_ERR_ARG,
                       "uIndx/uiIndx entry %lld of site %d does not name a "
                       "site holding it as a neighbor",
                       j, i);
    }
    for (int k = 0; k < lu[(size_t)n + i]; k++)
      if (nnIndx[lu[i] + k] >= i)
        return set_err(NNGP_ERR_ARG,
                       "site %d has neighbor %d that is not a predecessor", i,
                       nnIndx[lu[i] + k]);
  }
  if (tot != ntot)
    return set_err(NNGP_ERR_ARG,
                   "uIndxLU lists %lld entries, nnIndxLU %lld", tot, ntot);
  return NNGP_OK;
}

// Phases from a label per site: order lists the sites phase by phase
// (ascending inside a phase), phase[L] .. phase[L+1] is phase L.
static void phases_from_labels(int n, const std::vector<int> &lab, int nlab,
                               std::vector<int> &order,
                               std::vector<int> &phase) {
  phase.assign((size_t)nlab + 1, 0);
  for (int i = 0; i < n; i++) phase[(size_t)lab[i] + 1]++;
  for (int L = 0; L < nlab; L++) phase[(size_t)L + 1] += phase[L];
  order.assign((size_t)n, 0);
  std::vector<int> pos(phase.begin(), phase.end() - 1);
  for (int i = 0; i < n; i++) order[(size_t)pos[lab[i]]++] = i;
}

// Dependency levels of the sequential sweep: a site sits above its
// neighbors and above every earlier co-neighbor of each site holding it as
// a neighbor.  cm[j] is the highest level among j's neighbors placed so far
// (they are placed in ascending order).
static void sweep_levels(int n, const int *nnIndx, const int *lu,
                         const int *uIndx, const int *ulu,
                         std::vector<int> &order, std::vector<int> &phase) {
  std::vector<int> lev((size_t)n, 0), cm((size_t)n, -1);
  int nlev = 0;
  for (int i = 0; i < n; i++) {
    int l = -1;
    for (int k = 0; k < lu[(size_t)n + i]; k++)
      l = std::max(l, lev[nnIndx[lu[i] + k]]);
    const int uo = ulu[i], uc = ulu[(size_t)n + i];
    for (int j = 0; j < uc; j++) l = std::max(l, cm[uIndx[uo + j]]);
    lev[i] = l + 1;
    for (int j = 0; j < uc; j++) {
      int &c = cm[uIndx[uo + j]];
      c = std::max(c, lev[i]);
    }
    nlev = std::max(nlev, lev[i] + 1);
  }
  phases_from_labels(n, lev, nlev, order, phase);
}

// Greedy colouring of the moral graph in site order with 256-colour masks:
// a site avoids its neighbors' colours and the colours already taken among
// the neighbors of each site holding it as a neighbor.
constexpr int kMaxColours = 256;

static int sweep_colours(int n, const int *nnIndx, const int *lu,
                         const int *uIndx, const int *ulu,
                         std::vector<int> &order, std::vector<int> &phase) {
  constexpr int W = kMaxColours / 64;
  std::vector<uint64_t> mask((size_t)n * W, 0);
  std::vector<int> col((size_t)n, 0);
  int ncol = 0;
  for (int i = 0; i < n; i++) {
    uint64_t used[W] = {};
    for (int k = 0; k < lu[(size_t)n + i]; k++) {
      const int c = col[nnIndx[lu[i] + k]];
      used[c >> 6] |= 1ull << (c & 63);
    }
    const int uo = ulu[i], uc = ulu[(size_t)n + i];
    for (int j = 0; j < uc; j++) {
      const uint64_t *mj = &mask[(size_t)uIndx[uo + j] * W];
      for (int q = 0; q < W; q++) used[q] |= mj[q];
    }
    int c = -1;
    for (int q = 0; q < W && c < 0; q++)
      if (~used[q]) c = q * 64 + __builtin_ctzll(~used[q]);
    if (c < 0)
      return set_err(NNGP_ERR_UNSUPPORTED,
                     "the moral graph needs more than %d colours at site %d",
                     kMaxColours, i);
    col[i] = c;
    ncol = std::max(ncol, c + 1);
    for (int j = 0; j < uc; j++)
      mask[(size_t)uIndx[uo + j] * W + (c >> 6)] |= 1ull << (c & 63);
  }
  phases_from_labels(n, col, ncol, order, phase);
  return NNGP_OK;
}

// Standard exponential and Gamma deviates with R's methods, drawn from the
// sampler's source: Ahrens & Dieter's algorithm SA for the exponential
// (R's exp_rand) and their algorithm GD (Comm. ACM 25:47-54, 1982) for
// Gamma(a, scale), a >= 1 (R's rgamma), with exp(q) - 1 taken as expm1(q).
// The shape is fixed over a run, so its constants are set once.
struct GammaGD {
  double a, s2, s, d, q0, b, si, c;
  explicit GammaGD(double a_) : a(a_) {
    s2 = a - 0.5;
    s = std::sqrt(s2);
    d = 5.656854 - 12.0 * s;
    const double r = 1.0 / a;
    static const double qc[7] = {0.04166669, 0.02083148, 0.00801191,
                                 0.00144121, -7.388e-5,  2.4511e-4,
                                 2.424e-4};
    double q = qc[6];
    for (int k = 5; k >= 0; k--) q = q * r + qc[k];
    q0 = q * r;
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
  }
  // log of the quotient of the gamma density and the normal hat at t
  double quot(double t) const {
    const double v = t / (s + s);
    if (std::fabs(v) > 0.25)
      return q0 - s * t + 0.25 * t * t + (s2 + s2) * std::log(1.0 + v);
    static const double ac[7] = {0.3333333,  -0.250003, 0.2000062,
                                 -0.1662921, 0.1423657, -0.1367177,
                                 0.1233795};
    double pa = ac[6];
    for (int k = 5; k >= 0; k--) pa = pa * v + ac[k];
    return q0 + 0.5 * t * t * pa * v;
  }
  static double exp_rand(Rng &R) {
    // q[k] = sum_{j=1..k+1} (ln 2)^j / j!
    static const double q[16] = {
        0.6931471805599453, 0.933373687519046,  0.9888777961838676,
        0.9984959252914961, 0.9998292811061389, 0.9999833164100728,
        0.9999985691438767, 0.9999998906925558, 0.9999999924734159,
        0.9999999995283275, 0.9999999999728814, 0.9999999999985598,
        0.999999999999929,  0.9999999999999968, 0.9999999999999999, 1.0};
    double u;
    do {
      u = R.unif_rand();
    } while (u <= 0.0 || u >= 1.0);
    double base = 0.0;
    while ((u += u) <= 1.0) base += q[0];
    u -= 1.0;
    if (u <= q[0]) return base + u;
    double umin = R.unif_rand();
    int k = 0;
    do {
      const double us = R.unif_rand();
      umin = std::min(umin, us);
      k++;
    } while (u > q[k]);
    return base + umin * q[0];
  }
  double draw(Rng &R, double scale) const {
    double t = R.norm_rand();
    double x = s + 0.5 * t;
    const double xx = x * x;
    if (t >= 0.0) return scale * xx;           // immediate acceptance
    double u = R.unif_rand();
    if (d * u <= t * t * t) return scale * xx;  // squeeze
    if (x > 0.0 && std::log(1.0 - u) <= quot(t)) return scale * xx;
    for (;;) {  // double-exponential hat
      const double e = exp_rand(R);
      u = R.unif_rand();
      u = u + u - 1.0;
      t = u < 0.0 ? b - si * e : b + si * e;
      if (t < -0.71874483771719) continue;
      const double q = quot(t);
      if (q > 0.0 && c * std::fabs(u) <= std::expm1(q) * std::exp(e - 0.5 * t * t))
        break;
    }
    x = s + 0.5 * t;
    return scale * x * x;
  }
};

static double logit_ab(double t, double a, double b) {
  return std::log((t - a) / (b - t));
}
static double logit_inv(double z, double a, double b) {
  return b - (b - a) / (1 + std::exp(z));
}

}  // namespace nngp

using namespace nngp;

// ---------------------------------------------------------------------------
// mkUIndx (src/nn.cpp:25-56)
// ---------------------------------------------------------------------------
extern "C" int nngp_mkUIndx(int n, int m, const int *nnIndx, int *uIndx,
                            int *uIndxLU, int *uiIndx, const int *nnIndxLU,
                            int searchType) {
  clear_err();
  if (n < 2 || m < 1 || m >= n)
    return set_err(NNGP_ERR_ARG, "need n >= 2 and 1 <= n.neighbors < n");
  if (!uIndx || !uIndxLU || !uiIndx)
    return set_err(NNGP_ERR_ARG, "uIndx/uIndxLU/uiIndx is NULL");
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  NNGP_TRY(use_device(nullptr));
  const long long nIndx = n_indx(n, m);
  std::vector<int> rowstart((size_t)n + 1, 0);
  for (int r = 0; r < n; r++)
    rowstart[(size_t)r + 1] = rowstart[r] + nnIndxLU[(size_t)n + r];
  const int total = rowstart[n];
  if (total > nIndx)
    return set_err(NNGP_ERR_ARG, "nnIndxLU counts exceed nIndx");
  Stream st;
  NNGP_TRY(st.create());
  DevBuf dnn, dlu, drs, keys, vals, skeys, svals, owner, ucnt, uoff, du, dui,
      tmp;
  NNGP_TRY(dnn.alloc(sizeof(int) * (size_t)std::max(nIndx, 1LL)));
  NNGP_TRY(dlu.alloc(sizeof(int) * 2 * (size_t)n));
  NNGP_TRY(drs.alloc(sizeof(int) * ((size_t)n + 1)));
  const size_t te = (size_t)std::max(total, 1);
  NNGP_TRY(keys.alloc(sizeof(int) * te));
  NNGP_TRY(vals.alloc(sizeof(int) * te));
  NNGP_TRY(skeys.alloc(sizeof(int) * te));
  NNGP_TRY(svals.alloc(sizeof(int) * te));
  NNGP_TRY(owner.alloc(sizeof(int) * te));
  NNGP_TRY(du.alloc(sizeof(int) * te));
  NNGP_TRY(dui.alloc(sizeof(int) * te));
  NNGP_TRY(ucnt.alloc(sizeof(int) * (size_t)n));
  NNGP_TRY(uoff.alloc(sizeof(int) * (size_t)n));
  NNGP_HIP(hipMemcpyAsync(dnn.p, nnIndx, sizeof(int) * (size_t)nIndx,
                          hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(dlu.p, nnIndxLU, sizeof(int) * 2 * (size_t)n,
                          hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemcpyAsync(drs.p, rowstart.data(), sizeof(int) * ((size_t)n + 1),
                          hipMemcpyHostToDevice, st.s));
  NNGP_HIP(hipMemsetAsync(ucnt.p, 0, sizeof(int) * (size_t)n, st.s));
  const int gb = (n + kUThreads - 1) / kUThreads;
  hipLaunchKernelGGL(k_u_enumerate, dim3(gb), dim3(kUThreads), 0, st.s,
                     dnn.as<int>(), dlu.as<int>(), drs.as<int>(), n,
                     keys.as<int>(), vals.as<int>(), owner.as<int>(),
                     ucnt.as<int>());
  NNGP_HIP(hipGetLastError());
  int end_bit = 1;
  while (end_bit < 31 && (1 << end_bit) < n) end_bit++;
  size_t sort_bytes = 0, scan_bytes = 0;
  NNGP_HIP(hipcub::DeviceRadixSort::SortPairs(
      nullptr, sort_bytes, keys.as<int>(), skeys.as<int>(), vals.as<int>(),
      svals.as<int>(), total, 0, end_bit, st.s));
  NNGP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, ucnt.as<int>(),
                                            uoff.as<int>(), n, st.s));
  NNGP_TRY(tmp.alloc(std::max(sort_bytes, scan_bytes)));
  if (total > 0) {
    size_t b = tmp.bytes;
    NNGP_HIP(hipcub::DeviceRadixSort::SortPairs(
        tmp.p, b, keys.as<int>(), skeys.as<int>(), vals.as<int>(),
        svals.as<int>(), total, 0, end_bit, st.s));
  }
  {
    size_t b = tmp.bytes;
    NNGP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, b, ucnt.as<int>(),
                                              uoff.as<int>(), n, st.s));
  }
  if (total > 0) {
    hipLaunchKernelGGL(k_u_place, dim3((total + kUThreads - 1) / kUThreads),
                       dim3(kUThreads), 0, st.s, skeys.as<int>(),
                       svals.as<int>(), owner.as<int>(), drs.as<int>(),
                       uoff.as<int>(), ucnt.as<int>(), total,
                       searchType == 1 ? 1 : 0, du.as<int>(), dui.as<int>());
    NNGP_HIP(hipGetLastError());
    NNGP_HIP(hipMemcpyAsync(uIndx, du.p, sizeof(int) * (size_t)total,
                            hipMemcpyDeviceToHost, st.s));
    NNGP_HIP(hipMemcpyAsync(uiIndx, dui.p, sizeof(int) * (size_t)total,
                            hipMemcpyDeviceToHost, st.s));
  }
  NNGP_HIP(hipMemcpyAsync(uIndxLU, uoff.p, sizeof(int) * (size_t)n,
                          hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipMemcpyAsync(uIndxLU + n, ucnt.p, sizeof(int) * (size_t)n,
                          hipMemcpyDeviceToHost, st.s));
  NNGP_HIP(hipStreamSynchronize(st.s));
  return NNGP_OK;
}

// ---------------------------------------------------------------------------
// sNNGP (src/sNNGP.cpp:90-482)
// ---------------------------------------------------------------------------
extern "C" int nngp_sNNGP(
    const double *y, const double *X, int p, int n, int m,
    const double *coords, int covModel, const int *nnIndx,
    const int *nnIndxLU, const int *uIndx, const int *uIndxLU,
    const int *uiIndx, const double *sigmaSqIG, const double *tauSqIG,
    const double *phiUnif, const double *nuUnif, const double *betaStarting,
    double sigmaSqStarting, double tauSqStarting, double phiStarting,
    double nuStarting, double sigmaSqTuning, double tauSqTuning,
    double phiTuning, double nuTuning, int nSamples, int nThreads,
    int verbose, int nReport, const nngp_rng *rng, const nngp_opts *opts,
    double *pBetaSamples, double *pThetaSamples, double *pWSamples,
    int *nAccept) {
  (void)nThreads;
  (void)sigmaSqTuning;  // sigma.sq and tau.sq are Gibbs steps
  (void)tauSqTuning;
  clear_err();
  if (covModel < 0 || covModel > 3)
    return set_err(NNGP_ERR_COV_MODEL,
                   "c++ Rf_error: cov.model is not correctly specified");
  if (n < 2 || m < 1 || m >= n)
    return set_err(NNGP_ERR_ARG, "need n >= 2 and 1 <= n.neighbors < n");
  if (m > 64)
    return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  if (p < 1 || p > kMaxP)
    return set_err(NNGP_ERR_UNSUPPORTED, "p = %d covariates (1..%d supported)",
                   p, kMaxP);
  if (!y || !X || !coords || !sigmaSqIG || !tauSqIG || !phiUnif ||
      !betaStarting || !pBetaSamples || !pThetaSamples || !pWSamples ||
      nSamples < 0)
    return set_err(NNGP_ERR_ARG, "NULL buffer or negative nSamples");
  const bool matern = covModel == 2;
  if (matern && !nuUnif) return set_err(NNGP_ERR_ARG, "nuUnif is NULL");
  const int sweep = opts ? opts->sweep : NNGP_SWEEP_SITE_ORDER;
  if (sweep != NNGP_SWEEP_SITE_ORDER && sweep != NNGP_SWEEP_COLOURED)
    return set_err(NNGP_ERR_ARG, "unknown sweep order %d", sweep);
  const double sA = sigmaSqIG[0], sB = sigmaSqIG[1];
  const double tA = tauSqIG[0], tB = tauSqIG[1];
  if (!(sA + n / 2.0 >= 1.0) || !(tA + n / 2.0 >= 1.0))
    return set_err(NNGP_ERR_UNSUPPORTED,
                   "Gamma shape below 1 (IG shape + n/2 < 1)");
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  const long long nIndx = n_indx(n, m);
  NNGP_TRY(check_transpose(n, nnIndx, nnIndxLU, uIndx, uIndxLU, uiIndx, nIndx));
  NNGP_TRY(use_device(opts));

  static const char *names[] = {"exponential", "spherical", "matern",
                                "gaussian"};
  std::vector<int> order, phase;
  if (sweep == NNGP_SWEEP_SITE_ORDER)
    sweep_levels(n, nnIndx, nnIndxLU, uIndx, uIndxLU, order, phase);
  else
    NNGP_TRY(sweep_colours(n, nnIndx, nnIndxLU, uIndx, uIndxLU, order, phase));
  const int nph = (int)phase.size() - 1;
  if (verbose) {
    printf("----------------------------------------\n");
    printf("\tModel description\n");
    printf("----------------------------------------\n");
    printf("NNGP Latent model fit with %i observations.\n\n", n);
    printf("Number of covariates %i (including intercept if specified).\n\n",
           p);
    printf("Using the %s spatial correlation model.\n\n", names[covModel]);
    printf("Using %i nearest neighbors.\n\n", m);
    printf("Number of MCMC samples %i.\n\n", nSamples);
    printf("Priors and hyperpriors:\n\tbeta flat.\n");
    printf("\tsigma.sq IG hyperpriors shape=%.5f and scale=%.5f\n", sA, sB);
    printf("\ttau.sq IG hyperpriors shape=%.5f and scale=%.5f\n", tA, tB);
    printf("\tphi Unif hyperpriors a=%.5f and b=%.5f\n", phiUnif[0],
           phiUnif[1]);
    if (matern)
      printf("\tnu Unif hyperpriors a=%.5f and b=%.5f\n", nuUnif[0],
             nuUnif[1]);
    printf("\nComputed on HIP device %d; w sweep in %d %s.\n",
           opts ? opts->device : 0, nph,
           sweep == NNGP_SWEEP_SITE_ORDER ? "dependency levels (site order)"
                                          : "colour classes");
    fflush(stdout);
  }

  // device state: w is the model's one data column
  Model md;
  {
    std::vector<double> w0((size_t)n, 0.0);
    NNGP_TRY(md.create(n, m, 0, coords, nnIndx, nnIndxLU, nullptr, w0.data(),
                       covModel, opts));
  }
  hipStream_t st = md.st.s;
  DevBuf dy, dX, du, dui, duoc, dmeta, dR, dphase, dz, dpart, dsum, dwout,
      dbpos, dBt;
  NNGP_TRY(dy.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(dX.alloc(sizeof(double) * (size_t)n * p));
  NNGP_TRY(du.alloc(sizeof(int) * (size_t)std::max(nIndx, 1LL)));
  NNGP_TRY(dui.alloc(sizeof(int) * (size_t)std::max(nIndx, 1LL)));
  NNGP_TRY(duoc.alloc(sizeof(int2) * (size_t)std::max(nIndx, 1LL)));
  NNGP_TRY(dmeta.alloc(sizeof(int4) * (size_t)n));
  NNGP_TRY(dR.alloc(sizeof(double2) * (size_t)n));
  NNGP_TRY(dphase.alloc(sizeof(int) * phase.size()));
  NNGP_TRY(dpart.alloc(sizeof(double) * kRedMax * kLatStats));
  NNGP_TRY(dsum.alloc(sizeof(double) * kLatStats));
  NNGP_TRY(dwout.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(dz.alloc(sizeof(double) * (size_t)n));  // caller's or Philox
  NNGP_HIP(hipMemcpyAsync(dy.p, y, dy.bytes, hipMemcpyHostToDevice, st));
  NNGP_HIP(hipMemcpyAsync(dX.p, X, dX.bytes, hipMemcpyHostToDevice, st));
  long long ut = 0;
  for (int i = 0; i < n; i++)
    ut = std::max(ut, (long long)uIndxLU[i] + uIndxLU[(size_t)n + i]);
  std::vector<int2> huoc((size_t)std::max(ut, 1LL));
  for (long long t = 0; t < ut; t++) {
    const int jj = uIndx[t];
    // entries outside every site's range are never read; keep them valid
    const bool ok = jj >= 0 && jj < n;
    huoc[(size_t)t] = make_int2(ok ? nnIndxLU[jj] : 0,
                                ok ? nnIndxLU[(size_t)n + jj] : 0);
  }
  // position in B of each transpose entry's coefficient
  std::vector<int> hbpos((size_t)std::max(ut, 1LL), 0);
  for (long long t = 0; t < ut; t++)
    hbpos[(size_t)t] = huoc[(size_t)t].x + uiIndx[t];
  NNGP_TRY(dbpos.alloc(sizeof(int) * hbpos.size()));
  NNGP_TRY(dBt.alloc(sizeof(double) * hbpos.size()));
  NNGP_HIP(hipMemcpyAsync(dbpos.p, hbpos.data(), dbpos.bytes,
                          hipMemcpyHostToDevice, st));
  std::vector<int4> hmeta((size_t)n);
  for (int k = 0; k < n; k++) {
    const int i = order[(size_t)k];
    hmeta[(size_t)k] = make_int4(i, uIndxLU[i], uIndxLU[(size_t)n + i], 0);
  }
  if (ut > 0) {
    NNGP_HIP(hipMemcpyAsync(du.p, uIndx, sizeof(int) * (size_t)ut,
                            hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(dui.p, uiIndx, sizeof(int) * (size_t)ut,
                            hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(duoc.p, huoc.data(), sizeof(int2) * (size_t)ut,
                            hipMemcpyHostToDevice, st));
  }
  NNGP_HIP(hipMemcpyAsync(dmeta.p, hmeta.data(), dmeta.bytes,
                          hipMemcpyHostToDevice, st));
  NNGP_HIP(hipMemcpyAsync(dphase.p, phase.data(), dphase.bytes,
                          hipMemcpyHostToDevice, st));
  PinnedBuf hz, hsum, hw;
  if (rng) NNGP_TRY(hz.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(hsum.alloc(sizeof(double) * kLatStats));
  // w samples leave through pinned staging on a second stream: the copy of
  // iteration s and its move into the caller's array overlap iteration
  // s + 1's sweep
  NNGP_TRY(hw.alloc(sizeof(double) * (size_t)n));
  Stream st2;
  NNGP_TRY(st2.create());
  struct Ev {
    hipEvent_t e = nullptr;
    ~Ev() {
      if (e) (void)hipEventDestroy(e);
    }
  } evGather, evCopied;
  NNGP_HIP(hipEventCreateWithFlags(&evGather.e, hipEventDisableTiming));
  NNGP_HIP(hipEventCreateWithFlags(&evCopied.e, hipEventDisableTiming));
  struct Drain {  // on every return, before hw and the events go away
    hipStream_t s;
    ~Drain() { (void)hipStreamSynchronize(s); }
  } drain{st2.s};
  int pendingS = -1;  // iteration whose w is in, or on its way to, hw
  auto flush_w = [&]() -> int {
    if (pendingS < 0) return NNGP_OK;
    NNGP_HIP(hipEventSynchronize(evCopied.e));
    std::memcpy(pWSamples + (size_t)pendingS * n, hw.p,
                sizeof(double) * (size_t)n);
    pendingS = -1;
    return NNGP_OK;
  };

  // launch plan of one sweep: [ph0, ph1) per launch, grid or block
  struct Seg {
    int ph0, ph1;
    bool grid;
  };
  std::vector<Seg> segs;
  for (int L = 0; L < nph;) {
    if (phase[(size_t)L + 1] - phase[L] >= kGridPhase) {
      segs.push_back({L, L + 1, true});
      L++;
    } else {
      int e = L;
      while (e < nph && phase[(size_t)e + 1] - phase[e] < kGridPhase) e++;
      segs.push_back({L, e, false});
      L = e;
    }
  }

  // X^T X (src/sNNGP.cpp:244)
  std::vector<double> XtX((size_t)p * p);
  for (int j = 0; j < p; j++)
    for (int i = 0; i < p; i++) {
      double s = 0.0;
      for (int l = 0; l < n; l++)
        s += X[(size_t)i * n + l] * X[(size_t)j * n + l];
      XtX[(size_t)j * p + i] = s;
    }

  Rng R(rng, opts ? opts->seed : 1ull);
  const GammaGD gtau(tA + n / 2.0), gsig(sA + n / 2.0);
  const int nTheta = matern ? 4 : 3;
  double th[4] = {sigmaSqStarting, tauSqStarting, phiStarting,
                  matern ? nuStarting : 0.0};
  const double pA = phiUnif[0], pB = phiUnif[1];
  const double nA = matern ? nuUnif[0] : 0.0, nB = matern ? nuUnif[1] : 0.0;
  std::vector<double> beta(betaStarting, betaStarting + p);
  std::vector<double> tpp((size_t)p * p), tp(p), tp2(p);
  auto cov_of = [&](double s2, double phi, double nu) {
    return make_cov(covModel, s2, 0.0, phi, matern ? nu : 0.0);
  };
  const int red_blocks =
      std::min(kRedMax, std::max(1, (n + kRedBlock - 1) / kRedBlock));
  auto reduce = [&](int mode, int slot) -> int {
    LatSumArgs L;
    (void)slot;  // the sweep's slot: FR holds its F
    L.FR = dR.as<double2>();
    L.y = dy.as<double>();
    L.X = dX.as<double>();
    L.w = md.Z.as<double>();
    for (int k = 0; k < kMaxP; k++) L.beta[k] = k < p ? beta[k] : 0.0;
    L.tau2 = th[1];
    L.n = n;
    L.p = p;
    L.ws = md.zs;
    L.mode = mode;
    L.part = dpart.as<double>();
    hipLaunchKernelGGL(k_lat_partial, dim3(red_blocks), dim3(kRedBlock), 0,
                       st, L);
    NNGP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_lat_final, dim3(1), dim3(kRedBlock), 0, st,
                       (const double *)dpart.as<double>(), red_blocks,
                       dsum.as<double>());
    NNGP_HIP(hipGetLastError());
    NNGP_HIP(hipMemcpyAsync(hsum.p, dsum.p, sizeof(double) * kLatStats,
                            hipMemcpyDeviceToHost, st));
    NNGP_HIP(hipStreamSynchronize(st));
    return NNGP_OK;
  };

  int cur = 0, accept = 0, batchAccept = 0, status = 0;
  // B and F of slot cur were computed at sigma.sq = sigF (log det ldF).
  // Without a nugget B does not depend on sigma.sq and F is proportional to
  // it, so the current theta's pass of the phi step (src/sNNGP.cpp:343-360)
  // is a rescaling of these and of the sweep's maintained residual sums.
  PassOut p0;
  NNGP_TRY(md.pass(cov_of(th[0], th[2], th[3]), cur, true, &p0));
  double sigF = th[0], ldF = p0.logDet;
  int bt_slot = -1;  // the B slot that dBt holds (slot cur is written only
                     // before the loop; proposals go to 1 - cur)
  for (int s = 0; s < nSamples; s++) {
    // w | rest: n normals in site order (the reference draws them in its
    // sweep, which visits the sites in that order)
    if (rng) {
      double *z = hz.as<double>();
      for (int i = 0; i < n; i++) z[i] = R.rnorm(0.0, 1.0);
      NNGP_HIP(hipMemcpyAsync(dz.p, hz.p, sizeof(double) * (size_t)n,
                              hipMemcpyHostToDevice, st));
    }
    SweepArgs A;
    A.B = md.B[cur].as<double>();
    A.F = md.F[cur].as<double>();
    A.y = dy.as<double>();
    A.X = dX.as<double>();
    A.z = dz.as<double>();
    A.w = md.Z.as<double>();
    A.nnIndx = md.nnIndx.as<int>();
    A.uIndx = du.as<int>();
    A.uiIndx = dui.as<int>();
    A.phase = dphase.as<int>();
    A.meta = dmeta.as<int4>();
    A.uoc = duoc.as<int2>();
    A.Bt = dBt.as<double>();
    A.lu = md.lu.as<int>();
    A.FR = dR.as<double2>();
    for (int k = 0; k < kMaxP; k++) A.beta[k] = k < p ? beta[k] : 0.0;
    A.tau2 = th[1];
    A.seed = opts ? opts->seed : 1ull;
    A.iter = s;
    A.n = n;
    A.p = p;
    A.ws = md.zs;
    A.nph = nph;
    A.fscale = th[0] == sigF ? 1.0 : th[0] / sigF;
    if (bt_slot != cur && ut > 0) {
      hipLaunchKernelGGL(k_gather_bt,
                         dim3((unsigned)((ut + kSweepGrid - 1) / kSweepGrid)),
                         dim3(kSweepGrid), 0, st, dBt.as<double>(), A.B,
                         dbpos.as<int>(), ut);
      NNGP_HIP(hipGetLastError());
      bt_slot = cur;
    }
    if (!rng) {
      hipLaunchKernelGGL(k_philox_fill, dim3((n + kSweepGrid - 1) / kSweepGrid),
                         dim3(kSweepGrid), 0, st, dz.as<double>(), n, A.seed,
                         A.iter);
      NNGP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_resid_init, dim3((n + kSweepGrid - 1) / kSweepGrid),
                       dim3(kSweepGrid), 0, st, A);
    NNGP_HIP(hipGetLastError());
    for (const Seg &g : segs) {
      A.ph0 = g.ph0;
      A.ph1 = g.ph1;
      if (g.grid) {
        const int len = phase[(size_t)g.ph0 + 1] - phase[g.ph0];
        const long long thr = (long long)len * kGroup;
        hipLaunchKernelGGL(k_sweep_grid,
                           dim3((unsigned)((thr + kSweepGrid - 1) / kSweepGrid)),
                           dim3(kSweepGrid), 0, st, A);
      } else {
        hipLaunchKernelGGL(k_sweep_block, dim3(1), dim3(kSweepBlock), 0, st,
                           A);
      }
      NNGP_HIP(hipGetLastError());
    }
    NNGP_TRY(flush_w());  // the previous iteration's w, while the sweep runs
    // beta | rest (src/sNNGP.cpp:292-308) and the w quadratic form of the
    // current B, F for sigma.sq
    NNGP_TRY(reduce(0, cur));
    const double *hs = hsum.as<double>();
    for (int k = 0; k < p; k++) tp[k] = hs[k];
    const double aSig = hs[kMaxP];
    for (size_t k = 0; k < tpp.size(); k++) tpp[k] = XtX[k] / th[1];
    if (h_potrf_L(p, tpp.data()) || h_potri_L(p, tpp.data()))
      return set_err(NNGP_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
    h_symv_L(p, tpp.data(), tp.data(), tp2.data());
    if (h_potrf_L(p, tpp.data()))
      return set_err(NNGP_ERR_NOT_PD, "c++ Rf_error: dpotrf failed\n");
    for (int k = 0; k < p; k++) beta[k] = R.rnorm(0.0, 1.0);
    h_trmv_L(p, tpp.data(), beta.data());
    for (int k = 0; k < p; k++) beta[k] += tp2[k];
    // tau.sq | rest (:313-317)
    NNGP_TRY(reduce(1, cur));
    th[1] = 1.0 / gtau.draw(R, 1.0 / (tB + 0.5 * hsum.as<double>()[0]));
    // sigma.sq | rest (:322-341)
    const double sigOld = th[0];
    th[0] = 1.0 / gsig.draw(R, 1.0 / (sB + 0.5 * aSig * th[0]));
    // phi [, nu]: Metropolis with B, F of the new sigma.sq (:343-420); the
    // candidate is drawn before both passes, the draw order is unchanged
    const double phiC =
        logit_inv(R.rnorm(logit_ab(th[2], pA, pB), phiTuning), pA, pB);
    const double nuC =
        matern ? logit_inv(R.rnorm(logit_ab(th[3], nA, nB), nuTuning), nA, nB)
               : 0.0;
    // the current theta at the new sigma.sq: F = F(sigF) th[0] / sigF, so
    // log det = ldF + n log(th[0] / sigF); the quadratic form of the swept
    // residuals (aSig, taken with F at sigOld) scales by sigOld / th[0]
    const double ldCur = ldF + n * std::log(th[0] / sigF);
    const double qCur = aSig * (sigOld / th[0]);
    PassOut pn;
    NNGP_TRY(md.pass(cov_of(th[0], phiC, nuC), 1 - cur, true, &pn));
    double lpCur = -0.5 * ldCur - 0.5 * qCur;
    lpCur += std::log(th[2] - pA) + std::log(pB - th[2]);
    if (matern) lpCur += std::log(th[3] - nA) + std::log(nB - th[3]);
    double lpCand = -0.5 * pn.logDet - 0.5 * gram_at(pn.gram, 0, 0);
    lpCand += std::log(phiC - pA) + std::log(pB - phiC);
    if (matern) lpCand += std::log(nuC - nA) + std::log(nB - nuC);
    if (R.runif(0.0, 1.0) <= std::exp(lpCand - lpCur)) {
      cur = 1 - cur;
      th[2] = phiC;
      if (matern) th[3] = nuC;
      sigF = th[0];
      ldF = pn.logDet;
      accept++;
      batchAccept++;
    }
    for (int k = 0; k < p; k++) pBetaSamples[(size_t)s * p + k] = beta[k];
    for (int k = 0; k < nTheta; k++)
      pThetaSamples[(size_t)s * nTheta + k] = th[k];
    // w of this iteration: gathered on st, copied to hw on st2 (the
    // previous copy has been waited for in flush_w, so hw and dwout are
    // free), moved to pWSamples during the next iteration
    hipLaunchKernelGGL(k_gather_w, dim3((n + 255) / 256), dim3(256), 0, st,
                       (const double *)md.Z.as<double>(), md.zs, n,
                       dwout.as<double>());
    NNGP_HIP(hipGetLastError());
    NNGP_HIP(hipEventRecord(evGather.e, st));
    NNGP_HIP(hipStreamWaitEvent(st2.s, evGather.e, 0));
    NNGP_HIP(hipMemcpyAsync(hw.p, dwout.p, sizeof(double) * (size_t)n,
                            hipMemcpyDeviceToHost, st2.s));
    NNGP_HIP(hipEventRecord(evCopied.e, st2.s));
    pendingS = s;
    if (status == nReport) {
      if (verbose) {
        printf("Sampled: %i of %i, %3.2f%%\n", s, nSamples,
               100.0 * s / nSamples);
        printf("Report interval Metrop. Acceptance rate: %3.2f%%\n",
               100.0 * batchAccept / nReport);
        printf("Overall Metrop. Acceptance rate: %3.2f%%\n",
               100.0 * accept / s);
        printf("-------------------------------------------------\n");
        fflush(stdout);
      }
      batchAccept = 0;
      status = 0;
    }
    status++;
    if (nAccept) *nAccept = accept;
    if (opts && opts->interrupt && opts->interrupt(opts->interrupt_arg)) {
      NNGP_TRY(flush_w());
      return set_err(NNGP_ERR_INTERRUPT, "interrupted after %d samples", s + 1);
    }
  }
  NNGP_TRY(flush_w());
  if (verbose) {
    printf("Sampled: %i of %i, %3.2f%%\n", nSamples, nSamples, 100.0);
    printf("Overall Metrop. Acceptance rate: %3.2f%%\n",
           nSamples ? 100.0 * accept / nSamples : 0.0);
    fflush(stdout);
  }
  if (nAccept) *nAccept = accept;
  return NNGP_OK;
}
