// nngp_rep.hip — rNNGPReplicated (src/rNNGPReplicated.cpp:78-230): replicate
// data sets y.rep[, s] = v + X beta_s from the NNGP of posterior sample s,
// where v_i = B_i . v_N(i) + sqrt(F_i) z_i is drawn site by site in the NNGP
// order (:183-194) with B, F of theta_s (C = s2 R + t2 I, as in updateBF).
//
// Device design.
//  * B and F of a group of G samples come from G updateBF launches (the
//    same kernels as the samplers, without the sums) into per-sample
//    buffers, so nothing crosses PCIe but the normals and the replicates.
//  * The draw is sequential only along the neighbor DAG: site i waits for
//    its neighbors, all of which precede it.  Sites are grouped by level
//    (1 + the largest level among their neighbors, computed on the host
//    once per neighbor index); one block per sample walks the levels in
//    order with a barrier between them.  The depth grows like sqrt(n) for
//    x-ordered sites (about 8.5e3 levels at n = 1e6, m = 15), so one
//    sample's draw takes milliseconds, and the G samples of a group run on
//    G CUs at once.
//  * The normals: with a caller's generator (R's norm_rand through the
//    shim) they are drawn on the host in the reference's order, all n of
//    sample s before sample s + 1, so the replicates follow R's stream.
//    Without one they are counter-based Philox normals keyed by
//    (seed, site, sample), drawn in the kernel.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nngp_host.hpp"

namespace nngp {

constexpr int kRepThreads = 1024;

struct RepArgs {
  const double *B, *F, *z, *X, *beta;
  const int *nnIndx, *lu, *order, *loff;
  const int4 *sites;  // per position of order: {site, offset, count, 0}
  double *v, *y;
  long long bstride;  // doubles between samples' B
  int n, nlev, p, nS, s0;
  unsigned long long seed;  // Philox key when z == nullptr
};

// Block g: sample s0 + g.  Levels in order; within a level every site's
// neighbors are final, so its threads may run in any order.
__global__ __launch_bounds__(kRepThreads) void k_rep_draw(RepArgs A) {
  const int g = (int)blockIdx.x;
  const int n = A.n;
  const double *B = A.B + (size_t)g * A.bstride;
  const double *F = A.F + (size_t)g * n;
  const double *z = A.z ? A.z + (size_t)g * n : nullptr;
  double *v = A.v + (size_t)g * n;
  const int s = A.s0 + g;
  for (int L = 0; L < A.nlev; L++) {
    const int a = A.loff[L], b = A.loff[L + 1];
    for (int k = a + (int)threadIdx.x; k < b; k += (int)blockDim.x) {
      const int i = A.order[k];
      const int off = A.lu[i], cnt = A.lu[n + i];
      double acc = 0.0;
      for (int j = 0; j < cnt; j++) acc += B[off + j] * v[A.nnIndx[off + j]];
      const double zi =
          z ? z[i] : philox_normal(A.seed, (uint32_t)i, (uint32_t)s);
      v[i] = acc + sqrt(F[i]) * zi;
    }
    __syncthreads();
  }
  // y.rep = v + X beta_s (dgemv at src/rNNGPReplicated.cpp:196)
  for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) {
    double yi = v[i];
    for (int j = 0; j < A.p; j++)
      yi += A.X[(size_t)j * n + i] * A.beta[s + (size_t)j * A.nS];
    A.y[(size_t)g * n + i] = yi;
  }
}

// k_rep_draw for m <= M with one site a thread per level.  Slot t of
// level L is sites[loff[L] + t] = {site, its neighbor offset, count}.  The
// slot of level L + 2 and the rows of level L + 1 (neighbor ids, B row, F
// and the normal) are requested while level L waits for its neighbors' v,
// and nothing waits for them before that wait: each level costs one memory
// round trip and the barrier.  A level with more sites than threads runs
// its other sites as k_rep_draw does.  Every site's arithmetic is
// k_rep_draw's, in the same order.
constexpr int kRepPfThreads = 256;

template <int M>
struct RepSlot {
  int i;  // the site, or -1 when the level has no site for this thread
  int cnt;
  int nb[M];
  double b[M];
  double f, z;  // F_i and the normal
};

__device__ __forceinline__ int4 rep_site(const RepArgs &A, int L, int t) {
  int4 q = make_int4(-1, 0, 0, 0);
  if (L < A.nlev) {
    const int a = A.loff[L], b = A.loff[L + 1];
    if (a + t < b) q = A.sites[a + t];
  }
  return q;
}

// Straight-line code (no branch per slot, which would make the compiler
// wait for these loads where the branches join): slots past the count read
// slot 0 of the site (or of the index for an empty level slot, q = {-1, 0,
// 0, 0}) and are then set to 0.
template <int M>
__device__ __forceinline__ void rep_rows(const RepArgs &A, const double *B,
                                         const double *F, const double *z,
                                         int s, int4 q, RepSlot<M> &r) {
  r.i = q.x;
  r.cnt = q.z;
#pragma unroll
  for (int j = 0; j < M; j++) {
    const int jj = j < q.z ? j : 0;
    const int nb = A.nnIndx[q.y + jj];
    const double b = B[q.y + jj];
    r.nb[j] = j < q.z ? nb : 0;
    r.b[j] = j < q.z ? b : 0.0;
  }
  const int ii = q.x < 0 ? 0 : q.x;
  const double f = F[ii];
  const double zz =
      z ? z[ii] : philox_normal(A.seed, (uint32_t)ii, (uint32_t)s);
  r.f = q.x < 0 ? 0.0 : f;
  r.z = q.x < 0 ? 0.0 : zz;
}

template <int M>
__global__ __launch_bounds__(kRepPfThreads) void k_rep_draw_pf(RepArgs A) {
  const int g = (int)blockIdx.x;
  const int n = A.n;
  const int t = (int)threadIdx.x;
  const double *B = A.B + (size_t)g * A.bstride;
  const double *F = A.F + (size_t)g * n;
  const double *z = A.z ? A.z + (size_t)g * n : nullptr;
  double *v = A.v + (size_t)g * n;
  const int s = A.s0 + g;
  RepSlot<M> cur, nxt;
  int4 q2 = rep_site(A, 1, t);
  rep_rows<M>(A, B, F, z, s, rep_site(A, 0, t), cur);
  for (int L = 0; L < A.nlev; L++) {
    // this level's neighbor values first, every slot at once (padded slots
    // read v[0], which no product uses); the loads for the levels ahead go
    // out behind them, so that the wait for these leaves those in flight
    double vn[M];
#pragma unroll
    for (int j = 0; j < M; j++) vn[j] = v[cur.nb[j]];
    const int4 q3 = rep_site(A, L + 2, t);
    rep_rows<M>(A, B, F, z, s, q2, nxt);
    // past the count b = 0 and the value is replaced by 0, so those terms
    // add +0 to a sum that is never -0: the sum is k_rep_draw's bit for bit
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < M; j++)
      acc += cur.b[j] * (j < cur.cnt ? vn[j] : 0.0);
    const double vi = acc + sqrt(cur.f) * cur.z;
    if (cur.i >= 0) v[cur.i] = vi;
    const int a = A.loff[L], b = A.loff[L + 1];
    for (int k = a + t + (int)blockDim.x; k < b; k += (int)blockDim.x) {
      const int i = A.order[k];
      const int off = A.lu[i], cnt = A.lu[n + i];
      double acc = 0.0;
      for (int j = 0; j < cnt; j++) acc += B[off + j] * v[A.nnIndx[off + j]];
      const double zi =
          z ? z[i] : philox_normal(A.seed, (uint32_t)i, (uint32_t)s);
      v[i] = acc + sqrt(F[i]) * zi;
    }
    __syncthreads();
    cur = nxt;
    q2 = q3;
  }
  for (int i = t; i < n; i += (int)blockDim.x) {
    double yi = v[i];
    for (int j = 0; j < A.p; j++)
      yi += A.X[(size_t)j * n + i] * A.beta[s + (size_t)j * A.nS];
    A.y[(size_t)g * n + i] = yi;
  }
}

// One block per sample of the group: the prefetching form for m <= 16.
static void rep_draw_launch(const RepArgs &A, int groups, int m,
                            hipStream_t st) {
  if (m <= 16)
    hipLaunchKernelGGL(k_rep_draw_pf<16>, dim3(groups), dim3(kRepPfThreads),
                       0, st, A);
  else
    hipLaunchKernelGGL(k_rep_draw, dim3(groups), dim3(kRepThreads), 0, st,
                       A);
}

// Level schedule of the neighbor DAG: order[] lists the sites level by
// level (ascending site index within a level), loff[L] .. loff[L+1] is
// level L.  Fails if a neighbor does not precede its site.
static int build_levels(int n, const int *nnIndx, const int *lu,
                        std::vector<int> &order, std::vector<int> &loff) {
  std::vector<int> lev((size_t)n, 0);
  int nlev = 0;
  for (int i = 0; i < n; i++) {
    int l = 0;
    const int off = lu[i], cnt = lu[(size_t)n + i];
    for (int j = 0; j < cnt; j++) {
      const int id = nnIndx[off + j];
      if (id < 0 || id >= i)
        return set_err(NNGP_ERR_ARG,
                       "site %d has neighbor %d that is not a predecessor", i,
                       id);
      l = std::max(l, lev[id] + 1);
    }
    lev[i] = l;
    nlev = std::max(nlev, l + 1);
  }
  loff.assign((size_t)nlev + 1, 0);
  for (int i = 0; i < n; i++) loff[(size_t)lev[i] + 1]++;
  for (int L = 0; L < nlev; L++) loff[(size_t)L + 1] += loff[L];
  order.assign((size_t)n, 0);
  std::vector<int> pos(loff.begin(), loff.end() - 1);
  for (int i = 0; i < n; i++) order[(size_t)pos[lev[i]]++] = i;
  return NNGP_OK;
}

// sites[k] = {order[k], its neighbor offset, its count, 0}, on the device.
static int upload_sites(int n, const int *lu, const std::vector<int> &order,
                        DevBuf &dsites, hipStream_t st) {
  std::vector<int4> h((size_t)n);
  for (int k = 0; k < n; k++) {
    const int i = order[(size_t)k];
    h[(size_t)k] = make_int4(i, lu[i], lu[(size_t)n + i], 0);
  }
  NNGP_TRY(dsites.alloc(sizeof(int4) * (size_t)n));
  NNGP_HIP(hipMemcpyAsync(dsites.p, h.data(), dsites.bytes,
                          hipMemcpyHostToDevice, st));
  NNGP_HIP(hipStreamSynchronize(st));
  return NNGP_OK;
}

int RepDrawer::init(int n_, const int *nnIndx, const int *lu) {
  n = n_;
  m = 0;
  for (int i = 0; i < n; i++) m = std::max(m, lu[(size_t)n + i]);
  std::vector<int> order, loff;
  NNGP_TRY(build_levels(n, nnIndx, lu, order, loff));
  nlev = (int)loff.size() - 1;
  NNGP_TRY(dorder.alloc(sizeof(int) * (size_t)n));
  NNGP_TRY(dloff.alloc(sizeof(int) * loff.size()));
  NNGP_HIP(hipMemcpy(dorder.p, order.data(), dorder.bytes,
                     hipMemcpyHostToDevice));
  NNGP_HIP(hipMemcpy(dloff.p, loff.data(), sizeof(int) * loff.size(),
                     hipMemcpyHostToDevice));
  NNGP_TRY(upload_sites(n, lu, order, dsites, nullptr));
  NNGP_TRY(dz.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(dv.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(dy.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(hz.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(hv.alloc(sizeof(double) * (size_t)n));
  return NNGP_OK;
}

int RepDrawer::draw(const double *B, const double *F, const int *nnIndx,
                    const int *lu, hipStream_t st) {
  NNGP_HIP(hipMemcpyAsync(dz.p, hz.p, sizeof(double) * (size_t)n,
                          hipMemcpyHostToDevice, st));
  RepArgs A;
  A.B = B;
  A.F = F;
  A.z = dz.as<double>();
  A.X = nullptr;
  A.beta = nullptr;
  A.nnIndx = nnIndx;
  A.lu = lu;
  A.order = dorder.as<int>();
  A.loff = dloff.as<int>();
  A.sites = dsites.as<int4>();
  A.v = dv.as<double>();
  A.y = dy.as<double>();
  A.bstride = 0;
  A.n = n;
  A.nlev = nlev;
  A.p = 0;  // y = v
  A.nS = 1;
  A.s0 = 0;
  A.seed = 0;
  rep_draw_launch(A, 1, m, st);
  NNGP_HIP(hipGetLastError());
  NNGP_HIP(hipMemcpyAsync(hv.p, dy.p, sizeof(double) * (size_t)n,
                          hipMemcpyDeviceToHost, st));
  NNGP_HIP(hipStreamSynchronize(st));
  return NNGP_OK;
}

}  // namespace nngp

using namespace nngp;

extern "C" int nngp_rNNGPReplicated(
    const double *X, int p, int n, int m, const double *coords, int covModel,
    const int *nnIndx, const int *nnIndxLU, const double *beta,
    const double *theta, int nSamples, int verbose, int nReport,
    const nngp_rng *rng, const nngp_opts *opts, double *yRep) {
  clear_err();
  if (covModel < 0 || covModel > 3)
    return set_err(NNGP_ERR_COV_MODEL,
                   "c++ Rf_error: cov.model is not correctly specified");
  if (n < 2 || m < 1 || m >= n)
    return set_err(NNGP_ERR_ARG, "need n >= 2 and 1 <= n.neighbors < n");
  if (m > 64)
    return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  if (p < 0 || nSamples < 0 || !coords || !theta || (p > 0 && (!X || !beta)) ||
      (nSamples > 0 && !yRep))
    return set_err(NNGP_ERR_ARG, "NULL buffer or negative size");
  NNGP_TRY(check_neighbors(n, m, nnIndx, nnIndxLU));
  if (nSamples == 0) return NNGP_OK;
  NNGP_TRY(use_device(opts));
  std::vector<int> order, loff;
  NNGP_TRY(build_levels(n, nnIndx, nnIndxLU, order, loff));
  const int nlev = (int)loff.size() - 1;
  const int nTheta = covModel == 2 ? 4 : 3;

  Model md;  // coordinates and neighbor index resident; no X, y
  NNGP_TRY(md.create(n, m, 0, coords, nnIndx, nnIndxLU, nullptr, nullptr,
                     covModel, opts));
  hipStream_t st = md.st.s;
  // samples per group: B, F, v, y and z of each sample
  const long long bstride = (md.nIndx + 1) & ~1LL;
  const double per = 8.0 * ((double)bstride + 4.0 * n);
  // The draw takes about as long for one sample as for 256 (one block per
  // CU, each walking the levels), so a group is as large as half the free
  // device memory allows, up to 256 samples; at least 16 GB is assumed.
  double budget = 16e9;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
    budget = std::max(budget, 0.5 * (double)free_b);
  int G = (int)std::min<double>(budget / per, 256.0);
  G = std::max(1, std::min(G, nSamples));
  DevBuf dB, dF, dz, dv, dy, dX, dbeta, dorder, dloff, dsites;
  NNGP_TRY(dB.alloc(sizeof(double) * (size_t)bstride * G));
  NNGP_TRY(dF.alloc(sizeof(double) * (size_t)n * G));
  NNGP_TRY(dv.alloc(sizeof(double) * (size_t)n * G));
  NNGP_TRY(dy.alloc(sizeof(double) * (size_t)n * G));
  if (rng) NNGP_TRY(dz.alloc(sizeof(double) * (size_t)n * G));
  if (p > 0) {
    NNGP_TRY(dX.alloc(sizeof(double) * (size_t)n * p));
    NNGP_TRY(dbeta.alloc(sizeof(double) * (size_t)nSamples * p));
    NNGP_HIP(hipMemcpyAsync(dX.p, X, dX.bytes, hipMemcpyHostToDevice, st));
    NNGP_HIP(hipMemcpyAsync(dbeta.p, beta, dbeta.bytes, hipMemcpyHostToDevice,
                            st));
  }
  NNGP_TRY(dorder.alloc(sizeof(int) * (size_t)n));
  NNGP_TRY(dloff.alloc(sizeof(int) * loff.size()));
  NNGP_HIP(hipMemcpyAsync(dorder.p, order.data(), dorder.bytes,
                          hipMemcpyHostToDevice, st));
  NNGP_HIP(hipMemcpyAsync(dloff.p, loff.data(), sizeof(int) * loff.size(),
                          hipMemcpyHostToDevice, st));
  NNGP_TRY(upload_sites(n, nnIndxLU, order, dsites, st));
  PinnedBuf hz;
  if (rng) NNGP_TRY(hz.alloc(sizeof(double) * (size_t)n * G));
  Rng R(rng, opts ? opts->seed : 1);
  if (verbose)
    std::printf("Computing %d replicates of the NNGP response model: %d "
                "observations, %d neighbors, %d dependency levels.\n",
                nSamples, n, m, nlev);
  // With device normals, each group's replicates leave on a copy stream
  // while the next group computes: two output buffers, used in turn.  A
  // group's buffer is free again once its copy has returned, which is
  // before the group after next is launched.  With a caller's generator
  // the pinned normals of a group must have reached the device before the
  // next group's are written, so each group is copied (and its stream
  // drained) before the next one starts.
  const bool twobuf = nSamples > G && !rng;
  DevBuf dy2;
  Stream cs;
  if (twobuf) {
    NNGP_TRY(dy2.alloc(sizeof(double) * (size_t)n * G));
    NNGP_TRY(cs.create());
  }
  hipEvent_t done[2] = {nullptr, nullptr};
  for (hipEvent_t &e : done)
    NNGP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct EventPair {
    hipEvent_t *e;
    ~EventPair() {
      for (int k = 0; k < 2; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } done_guard{done};
  struct Pending {
    int s0 = -1, g = 0, buf = 0;
  } pend;
  int status = 0;
  // copy group `pend` out (after its kernels), then check its failure flag
  auto copy_out = [&]() -> int {
    if (pend.s0 < 0) return NNGP_OK;
    hipStream_t q = twobuf ? cs.s : st;
    const double *src = pend.buf ? dy2.as<double>() : dy.as<double>();
    int herr = INT_MAX;
    NNGP_HIP(hipStreamWaitEvent(q, done[pend.buf], 0));
    NNGP_HIP(hipMemcpyAsync(&herr, md.err.p, sizeof(int),
                            hipMemcpyDeviceToHost, q));
    NNGP_HIP(hipMemcpyAsync(yRep + (size_t)pend.s0 * n, src,
                            sizeof(double) * (size_t)n * pend.g,
                            hipMemcpyDeviceToHost, q));
    NNGP_HIP(hipStreamSynchronize(q));
    if (herr != INT_MAX) {
      (void)hipStreamSynchronize(st);  // nothing may run past the return
      return set_err(NNGP_ERR_NOT_PD,
                     "c++ Rf_error: dpotrf failed\n (site %d of a sample in "
                     "%d..%d)",
                     herr, pend.s0, pend.s0 + pend.g - 1);
    }
    status += pend.g;
    if (verbose && nReport > 0 && status >= nReport) {
      std::printf("Sampled: %d of %d, %3.2f%%\n", pend.s0 + pend.g,
                  nSamples, 100.0 * (pend.s0 + pend.g) / nSamples);
      std::fflush(stdout);
      status = 0;
    }
    pend.s0 = -1;
    return NNGP_OK;
  };
  int gi = 0;
  for (int s0 = 0; s0 < nSamples; s0 += G, gi++) {
    const int g = std::min(G, nSamples - s0);
    const int buf = twobuf ? gi % 2 : 0;
    for (int k = 0; k < g; k++) {
      const double *th = theta + (size_t)(s0 + k) * nTheta;
      const CovArgs c = make_cov(covModel, th[0], th[1], th[2],
                                 covModel == 2 ? th[3] : 0.0);
      BfArgs a;
      a.xy = md.xy.as<double2>();
      a.nnIndx = md.nnIndx.as<int>();
      a.lu = md.lu.as<int>();
      a.n = n;
      a.m = m;
      a.B = dB.as<double>() + (size_t)k * bstride;
      a.F = dF.as<double>() + (size_t)k * n;
      a.Z = nullptr;
      a.nz = 0;
      a.zs = 0;
      a.err = md.err.as<int>();
      a.rows = nullptr;
      a.rows_cap = 0;
      a.c = c;
      a.variant = cov_variant(c.cov, c.nu);
      a.lu_closed = md.lu_closed ? 1 : 0;
      BfConfig cfg = bf_default(m, a.variant, true);
      if (md.have_tuned && bf_config_ok(md.tuned_cfg, m, a.variant))
        cfg = md.tuned_cfg;
      const int rc = bf_launch(a, cfg, 1, st);
      if (rc)
        return set_err(NNGP_ERR_HIP, "updateBF launch failed (%d)", rc);
    }
    if (rng) {  // n normals per sample, sample after sample (:184)
      double *z = hz.as<double>();
      for (int k = 0; k < g; k++)
        for (int i = 0; i < n; i++) z[(size_t)k * n + i] = R.rnorm(0.0, 1.0);
      NNGP_HIP(hipMemcpyAsync(dz.p, hz.p, sizeof(double) * (size_t)n * g,
                              hipMemcpyHostToDevice, st));
    }
    RepArgs A;
    A.B = dB.as<double>();
    A.F = dF.as<double>();
    A.z = rng ? dz.as<double>() : nullptr;
    A.X = dX.as<double>();
    A.beta = dbeta.as<double>();
    A.nnIndx = md.nnIndx.as<int>();
    A.lu = md.lu.as<int>();
    A.order = dorder.as<int>();
    A.loff = dloff.as<int>();
    A.sites = dsites.as<int4>();
    A.v = dv.as<double>();
    A.y = buf ? dy2.as<double>() : dy.as<double>();
    A.bstride = bstride;
    A.n = n;
    A.nlev = nlev;
    A.p = p;
    A.nS = nSamples;
    A.s0 = s0;
    A.seed = opts ? opts->seed : 1;
    rep_draw_launch(A, g, m, st);
    NNGP_HIP(hipGetLastError());
    if (twobuf) NNGP_TRY(copy_out());  // the previous group, while this one
                                       // computes
    NNGP_HIP(hipEventRecord(done[buf], st));
    pend.s0 = s0;
    pend.g = g;
    pend.buf = buf;
    if (!twobuf) NNGP_TRY(copy_out());  // this group, before the next
  }
  NNGP_TRY(copy_out());
  if (verbose) {
    std::printf("Sampled: %d of %d, %3.2f%%\n", nSamples, nSamples, 100.0);
    std::fflush(stdout);
  }
  return NNGP_OK;
}
