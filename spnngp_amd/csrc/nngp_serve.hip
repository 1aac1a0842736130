// nngp_serve.hip — kriging server: a fitted response model resident in HBM
// answering rNNGPPredict requests (src/rNNGPPredict.cpp:25-222, with the
// nn2 neighbor query of R/generics.R:642) as they arrive.
//
// Continuous batching.  A request is admitted when the device arena (its
// coordinates, covariates, neighbor list and q x S results) and the pinned
// host arena (inputs staged in, results copied out) both have room.  On
// admission its inputs go up and its nn2 query runs on the admission
// stream.  A scheduler thread then packs sites of all admitted requests into
// kriging launches of at most batch_sites sites, shortest remaining request
// first (with aging), keeping `depth` launches in flight; each launch's
// results go back to the pinned arena on a copy stream while the next
// launch runs.  A request completes when its last sites are back on the
// host.  Its draws z(i, s) come from a counter-based generator keyed by the
// request's seed, so a result does not depend on which other requests
// shared its launches.
#include <time.h>

#include <algorithm>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <thread>

#include "nngp_host.hpp"
#include "nngp_strips.hpp"

namespace nngp {

constexpr int kMaxSegs = 256;           // runs per batched launch
constexpr double kAgeLimit = 0.25;      // s before a request jumps the queue

static double mono_now() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// First-fit allocator of byte ranges of an arena, 256-byte granules.
class RangeAlloc {
 public:
  void init(size_t cap) {
    cap_ = cap;
    used_ = peak_ = 0;
    free_.clear();
    free_[0] = cap;
  }
  size_t alloc(size_t bytes) {  // offset, or SIZE_MAX when nothing fits
    bytes = up256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return off;
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = up256(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
};

enum ReqState { kPending = 0, kActive = 1, kDone = 2 };

struct Req {
  long long ticket = 0;
  int q = 0;
  unsigned long long seed = 0;
  std::vector<double> in;  // interleaved coordinates (2q), then X0 (pq)
  double t_submit = 0, t_admit = 0, t_done = 0;
  size_t d_off = 0, d_bytes = 0, h_off = 0, h_bytes = 0;
  int *d_err = nullptr;
  double2 *d_xy0 = nullptr;
  double *d_X0 = nullptr;
  int *d_nn0 = nullptr;
  double *d_y0 = nullptr;
  int *h_err = nullptr;
  double *h_y0 = nullptr;
  hipEvent_t ready = nullptr;
  bool waited = false;   // the kernel stream already waits on `ready`
  int next = 0;          // sites handed to launches so far
  int copied = 0;        // sites whose results are back on the host
  int parts_out = 0;     // runs of this request in launches not retired
  bool d_held = false;   // owns its device arena range
  int state = kPending;
  int status = NNGP_OK;
  bool reported = false;
  bool in_done_q = false;             // finished, not yet reported
  std::list<long long>::iterator dq;  // its place in done_q
};

struct Part {
  Req *r;
  int first, count;
};

struct Slot {
  PinnedBuf hseg;
  DevBuf dseg;
  hipEvent_t ks = nullptr, kev = nullptr, oev = nullptr;  // ks..kev: kernel
  std::vector<Part> parts;
};

struct Server {
  int n = 0, p = 0, m = 0, S = 0, nTheta = 0, cov = 0, device = 0;
  int batch_sites = 16384, depth = 3;
  size_t dcap = 0, hcap = 0;
  DevBuf xy, X, y, beta, theta, covs, arena;
  PinnedBuf harena;
  DevStrips strips;
  Stream s_in, s_k, s_out;
  PredArgs base;
  Slot slots[8];             // the first `depth` are used
  std::deque<int> inflight;  // slot indices in launch order
  std::vector<int> free_slots;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  RangeAlloc dalloc, halloc;  // dalloc: scheduler thread only
  std::deque<Req *> pending;
  std::vector<Req *> active;  // scheduler thread only
  std::map<long long, Req *> reqs;
  // finished requests not yet handed out by a wait, in completion order
  // (a request leaves it when a wait reports it or when it is released)
  std::list<long long> done_q;
  long long next_ticket = 0;
  bool stop = false;
  std::thread worker;
  int worker_status = NNGP_OK;
  // statistics
  long long n_submitted = 0, n_completed = 0, n_launches = 0;
  double site_samples = 0.0;
  double kernel_ms = 0.0, kernel_site_samples = 0.0;  // retired launches

  void unlist_locked(Req *r) {
    if (r->in_done_q) {
      done_q.erase(r->dq);
      r->in_done_q = false;
    }
  }

  int create(const double *Xh, const double *yh, const double *coords, int n_,
             int p_, int m_, const double *betaS, const double *thetaS,
             int S_, int cov_, const nngp_server_opts *o);
  void run();
  int stage(Req *r);
  int launch(int slot_idx);
  void retire_locked();
  void complete_locked(Req *r, int status);
  size_t dev_bytes(int q) const {
    return 256 + up256(16 * (size_t)q) + up256(8 * (size_t)p * q) +
           up256(4 * (size_t)m * q) + up256(8 * (size_t)q * S);
  }
  size_t host_bytes(int q) const {
    return 256 + up256(8 * (size_t)q * S) + up256(16 * (size_t)q) +
           up256(8 * (size_t)p * q);
  }
  ~Server();
};

int Server::create(const double *Xh, const double *yh, const double *coords,
                   int n_, int p_, int m_, const double *betaS,
                   const double *thetaS, int S_, int cov_,
                   const nngp_server_opts *o) {
  n = n_;
  p = p_;
  m = m_;
  S = S_;
  cov = cov_;
  nTheta = cov == 2 ? 4 : 3;
  device = o ? o->device : 0;
  dcap = (o && o->device_bytes) ? (size_t)o->device_bytes : ((size_t)8 << 30);
  hcap = (o && o->host_bytes) ? (size_t)o->host_bytes : ((size_t)4 << 30);
  if (o && o->batch_sites > 0) batch_sites = o->batch_sites;
  if (o && o->depth > 0) depth = std::min(o->depth, 8);
  NNGP_HIP(hipSetDevice(device));
  NNGP_TRY(s_in.create());
  NNGP_TRY(s_k.create());
  NNGP_TRY(s_out.create());
  NNGP_TRY(upload_xy(coords, n, xy, s_in.s));
  NNGP_TRY(X.alloc(sizeof(double) * (size_t)n * p));
  NNGP_TRY(y.alloc(sizeof(double) * (size_t)n));
  NNGP_TRY(beta.alloc(sizeof(double) * (size_t)p * S));
  NNGP_TRY(theta.alloc(sizeof(double) * (size_t)nTheta * S));
  NNGP_HIP(hipMemcpyAsync(X.p, Xh, X.bytes, hipMemcpyHostToDevice, s_in.s));
  NNGP_HIP(hipMemcpyAsync(y.p, yh, y.bytes, hipMemcpyHostToDevice, s_in.s));
  NNGP_HIP(hipMemcpyAsync(beta.p, betaS, beta.bytes, hipMemcpyHostToDevice, s_in.s));
  NNGP_HIP(hipMemcpyAsync(theta.p, thetaS, theta.bytes, hipMemcpyHostToDevice,
                          s_in.s));
  const bool generic = cov == 2 || m > 16;
  std::vector<CovArgs> hc;
  if (generic) {
    hc.resize(S);
    for (int s = 0; s < S; s++) {
      const double *th = thetaS + (size_t)s * nTheta;
      hc[s] = make_cov(cov, th[0], th[1], th[2], cov == 2 ? th[3] : 0.0);
    }
    NNGP_TRY(covs.alloc(sizeof(CovArgs) * hc.size()));
    NNGP_HIP(hipMemcpyAsync(covs.p, hc.data(), covs.bytes, hipMemcpyHostToDevice,
                            s_in.s));
  }
  HostStrips hs;
  build_strips(n, coords, coords + n, m, hs);
  NNGP_TRY(strips.upload(hs, s_in.s));
  NNGP_TRY(arena.alloc(dcap));
  NNGP_TRY(harena.alloc(hcap));
  dalloc.init(dcap);
  halloc.init(hcap);
  for (int k = 0; k < depth; k++) {
    NNGP_TRY(slots[k].hseg.alloc(sizeof(PredSeg) * kMaxSegs));
    NNGP_TRY(slots[k].dseg.alloc(sizeof(PredSeg) * kMaxSegs));
    NNGP_HIP(hipEventCreate(&slots[k].ks));
    NNGP_HIP(hipEventCreate(&slots[k].kev));
    NNGP_HIP(hipEventCreateWithFlags(&slots[k].oev, hipEventDisableTiming));
    free_slots.push_back(depth - 1 - k);
  }
  NNGP_HIP(hipStreamSynchronize(s_in.s));
  base = PredArgs();
  base.xy = xy.as<double2>();
  base.X = X.as<double>();
  base.y = y.as<double>();
  base.n = n;
  base.p = p;
  base.m = m;
  base.beta = beta.as<double>();
  base.theta = theta.as<double>();
  base.S = S;
  base.nTheta = nTheta;
  base.cov = cov;
  base.unit = make_cov(cov, 1.0, 0.0, 1.0, 0.0);
  base.cov_s = generic ? covs.as<CovArgs>() : nullptr;
  worker = std::thread([this] { run(); });
  return NNGP_OK;
}

// Inputs to the device and the request's nn2 query (admission stream).
int Server::stage(Req *r) {
  const size_t q = (size_t)r->q;
  char *h = static_cast<char *>(harena.p) + r->h_off;
  r->h_err = reinterpret_cast<int *>(h);
  r->h_y0 = reinterpret_cast<double *>(h + 256);
  double2 *hxy = reinterpret_cast<double2 *>(h + 256 + up256(8 * q * S));
  double *hX0 = reinterpret_cast<double *>(reinterpret_cast<char *>(hxy) +
                                           up256(16 * q));
  std::memcpy(hxy, r->in.data(), 16 * q);
  std::memcpy(hX0, r->in.data() + 2 * q, 8 * (size_t)p * q);
  std::vector<double>().swap(r->in);
  char *d = static_cast<char *>(arena.p) + r->d_off;
  r->d_err = reinterpret_cast<int *>(d);
  r->d_xy0 = reinterpret_cast<double2 *>(d + 256);
  r->d_X0 = reinterpret_cast<double *>(reinterpret_cast<char *>(r->d_xy0) +
                                       up256(16 * q));
  r->d_nn0 = reinterpret_cast<int *>(reinterpret_cast<char *>(r->d_X0) +
                                     up256(8 * (size_t)p * q));
  r->d_y0 = reinterpret_cast<double *>(reinterpret_cast<char *>(r->d_nn0) +
                                       up256(4 * (size_t)m * q));
  NNGP_HIP(hipMemcpyAsync(r->d_xy0, hxy, 16 * q, hipMemcpyHostToDevice, s_in.s));
  NNGP_HIP(hipMemcpyAsync(r->d_X0, hX0, 8 * (size_t)p * q, hipMemcpyHostToDevice,
                          s_in.s));
  if (fill_int_launch(r->d_err, INT_MAX, s_in.s))
    return set_err(NNGP_ERR_HIP, "flag reset kernel failed");
  Nn2Args a;
  a.s = strips.idx;
  a.q = r->d_xy0;
  a.nq = r->q;
  a.k = m;
  a.idx = r->d_nn0;
  a.dist = nullptr;
  if (nn2_launch(a, s_in.s)) return set_err(NNGP_ERR_HIP, "nn2 launch failed");
  NNGP_HIP(hipEventCreateWithFlags(&r->ready, hipEventDisableTiming));
  NNGP_HIP(hipEventRecord(r->ready, s_in.s));
  return NNGP_OK;
}

// One batched kriging launch over the parts in slots[k], then the copies of
// their results (copy stream, after the launch).
int Server::launch(int k) {
  Slot &sl = slots[k];
  PredSeg *hs = sl.hseg.as<PredSeg>();
  int blocks = 0;
  for (size_t j = 0; j < sl.parts.size(); j++) {
    const Part &pt = sl.parts[j];
    Req *r = pt.r;
    PredSeg g;
    g.xy0 = r->d_xy0;
    g.X0 = r->d_X0;
    g.nn0 = r->d_nn0;
    g.y0 = r->d_y0;
    g.err = r->d_err;
    g.seed = r->seed;
    g.q = r->q;
    g.first = pt.first;
    g.block0 = blocks;
    g.count = pt.count;
    hs[j] = g;
    blocks += pt.count;
    if (!r->waited) {
      NNGP_HIP(hipStreamWaitEvent(s_k.s, r->ready, 0));
      r->waited = true;
    }
  }
  NNGP_HIP(hipMemcpyAsync(sl.dseg.p, hs, sizeof(PredSeg) * sl.parts.size(),
                          hipMemcpyHostToDevice, s_k.s));
  PredArgs a = base;
  a.segs = sl.dseg.as<PredSeg>();
  a.nseg = (int)sl.parts.size();
  a.site0 = 0;
  a.nsite = blocks;
  NNGP_HIP(hipEventRecord(sl.ks, s_k.s));
  if (pred_launch(a, s_k.s)) return set_err(NNGP_ERR_HIP, "kriging launch failed");
  NNGP_HIP(hipEventRecord(sl.kev, s_k.s));
  NNGP_HIP(hipStreamWaitEvent(s_out.s, sl.kev, 0));
  for (const Part &pt : sl.parts) {
    Req *r = pt.r;
    const size_t pitch = sizeof(double) * (size_t)r->q;
    NNGP_HIP(hipMemcpy2DAsync(r->h_y0 + pt.first, pitch, r->d_y0 + pt.first,
                              pitch, sizeof(double) * (size_t)pt.count, S,
                              hipMemcpyDeviceToHost, s_out.s));
    if (pt.first + pt.count == r->q)
      NNGP_HIP(hipMemcpyAsync(r->h_err, r->d_err, sizeof(int),
                              hipMemcpyDeviceToHost, s_out.s));
  }
  NNGP_HIP(hipEventRecord(sl.oev, s_out.s));
  n_launches++;
  return NNGP_OK;
}

void Server::complete_locked(Req *r, int status) {
  r->t_done = mono_now();
  r->status = status;
  r->state = kDone;
  if (r->d_held && r->parts_out == 0) {  // else released when they retire
    dalloc.release(r->d_off, r->d_bytes);
    r->d_held = false;
  }
  active.erase(std::remove(active.begin(), active.end(), r), active.end());
  r->dq = done_q.insert(done_q.end(), r->ticket);
  r->in_done_q = true;
  n_completed++;
  if (status == NNGP_OK) site_samples += (double)r->q * S;
  cv_done.notify_all();
}

void Server::retire_locked() {
  while (!inflight.empty()) {
    const int k = inflight.front();
    const hipError_t e = hipEventQuery(slots[k].oev);
    if (e == hipErrorNotReady) break;
    inflight.pop_front();
    if (e == hipSuccess) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, slots[k].ks, slots[k].kev) == hipSuccess)
        kernel_ms += ms;
      long long sites = 0;
      for (const Part &pt : slots[k].parts) sites += pt.count;
      kernel_site_samples += (double)sites * S;
    }
    for (const Part &pt : slots[k].parts) {
      Req *r = pt.r;
      r->copied += pt.count;
      r->parts_out--;
      if (r->state != kDone) {
        if (e != hipSuccess)
          complete_locked(r, NNGP_ERR_HIP);
        else if (r->copied == r->q)
          complete_locked(r, *r->h_err == INT_MAX ? NNGP_OK : NNGP_ERR_NOT_PD);
      } else if (r->d_held && r->parts_out == 0) {
        dalloc.release(r->d_off, r->d_bytes);
        r->d_held = false;
      }
    }
    slots[k].parts.clear();
    free_slots.push_back(k);
  }
}

void Server::run() {
  if (hipSetDevice(device) != hipSuccess) worker_status = NNGP_ERR_HIP;
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    retire_locked();
    // admission, in arrival order, while both arenas have room
    while (!pending.empty()) {
      Req *r = pending.front();
      const size_t ho = halloc.alloc(r->h_bytes);
      if (ho == SIZE_MAX) break;
      const size_t dof = dalloc.alloc(r->d_bytes);
      if (dof == SIZE_MAX) {
        halloc.release(ho, r->h_bytes);
        break;
      }
      pending.pop_front();
      r->h_off = ho;
      r->d_off = dof;
      r->d_held = true;
      r->state = kActive;
      r->t_admit = mono_now();
      active.push_back(r);
      lk.unlock();
      const int st = stage(r);
      lk.lock();
      if (st != NNGP_OK) complete_locked(r, st);
    }
    // launches: shortest remaining request first; a request older than
    // kAgeLimit goes first regardless
    while (!free_slots.empty()) {
      std::vector<Req *> cand;
      for (Req *r : active)
        if (r->next < r->q) cand.push_back(r);
      if (cand.empty()) break;
      const double now = mono_now();
      std::sort(cand.begin(), cand.end(), [&](const Req *a, const Req *b) {
        const bool oa = now - a->t_submit > kAgeLimit;
        const bool ob = now - b->t_submit > kAgeLimit;
        if (oa != ob) return oa;
        if (oa) return a->ticket < b->ticket;
        const int ra = a->q - a->next, rb = b->q - b->next;
        return ra != rb ? ra < rb : a->ticket < b->ticket;
      });
      const int k = free_slots.back();
      free_slots.pop_back();
      Slot &sl = slots[k];
      int cap = batch_sites;
      for (Req *r : cand) {
        if (cap == 0 || (int)sl.parts.size() == kMaxSegs) break;
        const int c = std::min(cap, r->q - r->next);
        sl.parts.push_back({r, r->next, c});
        r->parts_out++;
        r->next += c;
        cap -= c;
      }
      lk.unlock();
      const int st = launch(k);
      lk.lock();
      if (st == NNGP_OK) {
        inflight.push_back(k);
      } else {  // nothing of this launch is queued beyond its failed call
        (void)hipStreamSynchronize(s_k.s);
        (void)hipStreamSynchronize(s_out.s);
        worker_status = st;
        for (const Part &pt : sl.parts) {
          pt.r->parts_out--;
          if (pt.r->state != kDone) complete_locked(pt.r, st);
        }
        sl.parts.clear();
        free_slots.push_back(k);
      }
    }
    if (stop && pending.empty() && active.empty() && inflight.empty()) break;
    if (!inflight.empty() || !active.empty()) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    } else {
      cv_work.wait(lk, [&] { return stop || !pending.empty(); });
    }
  }
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
    for (Req *r : pending) complete_locked(r, NNGP_ERR_INTERRUPT);
    pending.clear();
  }
  cv_work.notify_all();
  if (worker.joinable()) worker.join();
  for (auto &kv : reqs) {
    if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
    delete kv.second;
  }
  for (Slot &sl : slots) {
    if (sl.ks) (void)hipEventDestroy(sl.ks);
    if (sl.kev) (void)hipEventDestroy(sl.kev);
    if (sl.oev) (void)hipEventDestroy(sl.oev);
  }
}

}  // namespace nngp

using namespace nngp;

struct nngp_server {
  Server sv;
};

int nngp_server_create(nngp_server **out, const double *X, const double *y,
                       const double *coords, int n, int p, int m,
                       const double *betaSamples, const double *thetaSamples,
                       int nSamples, int covModel,
                       const nngp_server_opts *opts) {
  clear_err();
  if (!out) return set_err(NNGP_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (covModel < 0 || covModel > 3)
    return set_err(NNGP_ERR_COV_MODEL,
                   "c++ Rf_error: cov.model is not correctly specified");
  if (n < 1 || m < 1 || m > n)
    return set_err(NNGP_ERR_ARG, "need 1 <= n.neighbors <= n");
  if (m > 64) return set_err(NNGP_ERR_UNSUPPORTED, "n.neighbors = %d exceeds 64", m);
  if (p < 1 || p > kMaxP)
    return set_err(NNGP_ERR_UNSUPPORTED, "p = %d covariates (1..%d supported)",
                   p, kMaxP);
  if (nSamples < 1) return set_err(NNGP_ERR_ARG, "nSamples must be >= 1");
  if (!X || !y || !coords || !betaSamples || !thetaSamples)
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  if (opts && (opts->device < 0 || opts->device >= nngp_device_count()))
    return set_err(NNGP_ERR_ARG, "device %d not available", opts->device);
  if (nngp_device_count() < 1)
    return set_err(NNGP_ERR_HIP, "no HIP device visible");
  nngp_server *s = new (std::nothrow) nngp_server;
  if (!s) return set_err(NNGP_ERR_NOMEM, "host allocation failed");
  const int st = s->sv.create(X, y, coords, n, p, m, betaSamples, thetaSamples,
                              nSamples, covModel, opts);
  if (st != NNGP_OK) {
    delete s;
    return st;
  }
  *out = s;
  return NNGP_OK;
}

void nngp_server_destroy(nngp_server *srv) { delete srv; }

int nngp_server_submit(nngp_server *srv, int q, const double *X0,
                       const double *coords0, unsigned long long seed,
                       long long *ticket) {
  clear_err();
  if (!srv || !ticket) return set_err(NNGP_ERR_ARG, "NULL argument");
  Server &sv = srv->sv;
  if (q < 1) return set_err(NNGP_ERR_ARG, "q must be >= 1");
  if (!X0 || !coords0) return set_err(NNGP_ERR_ARG, "NULL buffer");
  Req *r = new (std::nothrow) Req;
  if (!r) return set_err(NNGP_ERR_NOMEM, "host allocation failed");
  r->q = q;
  r->seed = seed;
  r->d_bytes = sv.dev_bytes(q);
  r->h_bytes = sv.host_bytes(q);
  if (r->d_bytes > sv.dcap || r->h_bytes > sv.hcap) {
    delete r;
    return set_err(NNGP_ERR_NOMEM,
                   "request of %d sites needs %zu device / %zu host bytes; the "
                   "server arenas hold %zu / %zu", q, sv.dev_bytes(q),
                   sv.host_bytes(q), sv.dcap, sv.hcap);
  }
  r->in.resize((size_t)(2 + sv.p) * q);
  for (int i = 0; i < q; i++) {
    r->in[2 * (size_t)i] = coords0[i];
    r->in[2 * (size_t)i + 1] = coords0[(size_t)q + i];
  }
  std::memcpy(r->in.data() + 2 * (size_t)q, X0, sizeof(double) * (size_t)sv.p * q);
  r->t_submit = mono_now();
  {
    std::lock_guard<std::mutex> lk(sv.mu);
    if (sv.worker_status != NNGP_OK) {
      delete r;
      return set_err(sv.worker_status, "the server's scheduler stopped on an error");
    }
    r->ticket = sv.next_ticket++;
    sv.reqs[r->ticket] = r;
    sv.pending.push_back(r);
    sv.n_submitted++;
  }
  sv.cv_work.notify_one();
  *ticket = r->ticket;
  return NNGP_OK;
}

int nngp_server_wait(nngp_server *srv, long long *ticket, double timeout_s,
                     nngp_request_info *info) {
  clear_err();
  if (!srv || !ticket) return set_err(NNGP_ERR_ARG, "NULL argument");
  Server &sv = srv->sv;
  std::unique_lock<std::mutex> lk(sv.mu);
  const long long want = *ticket;
  if (want >= 0 && !sv.reqs.count(want))
    return set_err(NNGP_ERR_ARG, "unknown ticket %lld", want);
  auto ready = [&] {
    if (want >= 0) {
      auto it = sv.reqs.find(want);
      return it == sv.reqs.end() || it->second->state == kDone;
    }
    return !sv.done_q.empty();
  };
  if (timeout_s < 0) {
    sv.cv_done.wait(lk, ready);
  } else if (!sv.cv_done.wait_for(
                 lk, std::chrono::duration<double>(timeout_s), ready)) {
    return set_err(NNGP_ERR_TIMEOUT, "no request finished within %g s", timeout_s);
  }
  Req *r = nullptr;
  if (want >= 0) {
    auto it = sv.reqs.find(want);
    if (it == sv.reqs.end())
      return set_err(NNGP_ERR_ARG, "ticket %lld was released meanwhile", want);
    r = it->second;
  } else {
    r = sv.reqs.at(sv.done_q.front());
  }
  sv.unlist_locked(r);
  r->reported = true;
  *ticket = r->ticket;
  if (info) {
    info->q = r->q;
    info->nSamples = sv.S;
    info->status = r->status;
    info->t_submit = r->t_submit;
    info->t_admit = r->t_admit;
    info->t_done = r->t_done;
    info->y0 = r->status == NNGP_OK ? r->h_y0 : nullptr;
  }
  return NNGP_OK;
}

int nngp_server_release(nngp_server *srv, long long ticket) {
  clear_err();
  if (!srv) return set_err(NNGP_ERR_ARG, "NULL argument");
  Server &sv = srv->sv;
  std::lock_guard<std::mutex> lk(sv.mu);
  auto it = sv.reqs.find(ticket);
  if (it == sv.reqs.end()) return set_err(NNGP_ERR_ARG, "unknown ticket %lld", ticket);
  Req *r = it->second;
  if (r->state != kDone || r->parts_out > 0)
    return set_err(NNGP_ERR_ARG, "request %lld is still running", ticket);
  sv.unlist_locked(r);
  if (r->h_err) sv.halloc.release(r->h_off, r->h_bytes);
  if (r->ready) (void)hipEventDestroy(r->ready);
  sv.reqs.erase(it);
  delete r;
  return NNGP_OK;
}

int nngp_server_predict(nngp_server *srv, int q, const double *X0,
                        const double *coords0, unsigned long long seed,
                        double *pY0) {
  if (!pY0) {
    clear_err();
    return set_err(NNGP_ERR_ARG, "NULL buffer");
  }
  long long t = -1;
  NNGP_TRY(nngp_server_submit(srv, q, X0, coords0, seed, &t));
  nngp_request_info info;
  NNGP_TRY(nngp_server_wait(srv, &t, -1.0, &info));
  if (info.status == NNGP_OK)
    std::memcpy(pY0, info.y0, sizeof(double) * (size_t)q * info.nSamples);
  const int st = info.status;
  NNGP_TRY(nngp_server_release(srv, t));
  if (st != NNGP_OK)
    return set_err(st, st == NNGP_ERR_NOT_PD
                           ? "c++ Rf_error: dpotrf failed\n(a kriging system "
                             "of the request is not positive definite)"
                           : "request failed");
  return NNGP_OK;
}

int nngp_server_get_stats(nngp_server *srv, nngp_server_stats *st) {
  clear_err();
  if (!srv || !st) return set_err(NNGP_ERR_ARG, "NULL argument");
  Server &sv = srv->sv;
  std::lock_guard<std::mutex> lk(sv.mu);
  st->submitted = sv.n_submitted;
  st->completed = sv.n_completed;
  st->launches = sv.n_launches;
  st->pending = (int)sv.pending.size();
  st->active = (int)sv.active.size();
  st->site_samples = sv.site_samples;
  st->device_bytes_used = sv.dalloc.used();
  st->device_bytes_peak = sv.dalloc.peak();
  st->host_bytes_used = sv.halloc.used();
  st->host_bytes_peak = sv.halloc.peak();
  st->kernel_ms = sv.kernel_ms;
  st->kernel_site_samples = sv.kernel_site_samples;
  st->done_unreported = (long long)sv.done_q.size();
  return NNGP_OK;
}
