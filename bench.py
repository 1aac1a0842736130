#!/usr/bin/env python3
"""Headline benchmark: response-NNGP MCMC iterations per second on one GPU.

Workload (the at-scale response configuration): n = 1e6 sites ~ U[0,1]^2
ordered by x, X = [1, N(0,1), N(0,1)], beta = (1, 2, -1), w drawn from the
NNGP prior (exponential, sigma.sq = 2, phi = 12, no nugget, m = 15,
w_i = B_i w_N(i) + sqrt(F_i) z_i), tau.sq = 0.2.  A step is one
Metropolis-within-Gibbs iteration of rNNGP (beta draw, Q(r, r) at the
current theta, updateBF of the proposal with its log-det and quadratic
forms, accept/reject).  Data, neighbor index and B/F buffers are resident
in HBM before the timed region starts; the host keeps the sampler.

Prints one JSON line.  --gpus > 1 is refused: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MCMC iter/s @ n=1e6,m=15 (FP64 vs mixed) + kriging p50/p99 latency"
PEAK_FP64_TFLOPS = 78.6
# The chains time every 4th updateBF launch (and its sum launch) with HIP
# events: timing each one adds two event packets per kernel to the stream and
# slows the chain by about 5%; one in four costs about 1%.
PROF_EVERY = 4
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the dominant kernels, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) of this script, each scaled by what the
# counter reads for a copy kernel of known size: the committed summary's
# table rows "| key | kernel | workload | bytes |".
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_per_launch.md")


def traffic(key, kernel):
    """(bytes, note) of the summary row `key` that measured `kernel` (the
    launch this run used), else (None, why)."""
    seen = []
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                cells = [x.strip() for x in line.strip().strip("|").split("|")]
                if len(cells) >= 4 and cells[0] == key:
                    if cells[1].strip("`") == kernel:
                        return int(cells[3].replace(",", "")), \
                            f"profiles/hbm_traffic_per_launch.md row {key}"
                    seen.append(cells[1])
    except OSError:
        return None, "no traffic summary"
    if seen:
        return None, (f"summary rows {key} measured {', '.join(seen)}; "
                      f"this run used {kernel}")
    return None, f"no summary row {key}"


def f_krig(m, p):
    """Algorithmic flops of one (new site, posterior sample) of kriging: the
    m x m covariance from the site's shared distances (2 flops and one exp
    per pair), Cholesky m^3/3, two solves 2m^2, w.c and the mean and
    variance terms 2pm + 2p + 4m + 3, and the one sqrt."""
    pairs = m * (m + 1) / 2
    return m ** 3 / 3 + 2 * m * m + 2 * pairs + 2 * p * m + 2 * p + 4 * m \
        + 3 + pairs + 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def f_alg(m):
    """Algorithmic flops of one site's updateBF with m neighbors: Cholesky
    m^3/3, two triangular solves 2m^2, 7 flops per pair for the distance and
    kernel, 2m + 2 for F, and one per exp and per sqrt (m(m+1)/2 each)."""
    pairs = m * (m + 1) / 2
    return m ** 3 / 3 + 2 * m * m + 7 * pairs + 2 * m + 2 + 2 * pairs


def make_data(n, m, seed):
    import spnngp_amd as nn
    rng = np.random.default_rng(seed)
    c = rng.uniform(size=(n, 2))
    c = c[np.argsort(c[:, 0], kind="stable")]
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    beta = np.array([1.0, 2.0, -1.0])
    t0 = time.perf_counter()
    idx = nn.mkNNIndxCB(c, m)
    knn_s = time.perf_counter() - t0
    # w from the NNGP prior itself, exponential, sigma.sq = 2, phi = 12
    B, F, _ = nn.updateBF(c, idx["nnIndx"], idx["nnIndxLU"], m,
                          [2.0, 0.0, 12.0], "exponential")
    w = nn.prior_draw(B, F, idx["nnIndx"], idx["nnIndxLU"],
                      rng.normal(size=n))
    y = X @ beta + w + rng.normal(scale=np.sqrt(0.2), size=n)
    return c, X, y, idx, knn_s


def loop_rate(tm, iters):
    """Iterations per second of the sampler's loop, from the library's own
    marks: the first proposal to the last decision, the model's stream
    drained at both (the call's initial pass at the starting theta and its
    set-up are outside, as a continuing chain does not repeat them)."""
    if tm["loop_iters"] != iters or tm["loop_ms"] <= 0:
        raise RuntimeError("loop timing covers %d iterations, expected %d"
                           % (tm["loop_iters"], iters))
    return iters / (tm["loop_ms"] * 1e-3), tm["loop_ms"] * 1e-3


def mcmc_bench(c, X, y, idx, m, steps, warmup, start, precision):
    """Warm up, autotune the updateBF launch shape at the chain's current
    theta, then time `steps` MH iterations with the data resident.  The rate
    is the sampler loop's (`loop_rate`); `call_iters_s` is the whole call's,
    which adds the initial pass and the call's set-up and copy-out."""
    import spnngp_amd as nn
    md = nn.Model(c, idx["nnIndx"], idx["nnIndxLU"], m, "exponential", X, y,
                  precision=precision)
    w = md.rNNGP(nSamples=warmup, seed=11, **start)
    th_last = w["p.theta.samples"][:, -1]
    start_t = dict(start, sigmaSqStarting=th_last[0], tauSqStarting=th_last[1],
                   phiStarting=th_last[2],
                   betaStarting=w["p.beta.samples"][:, -1])
    cand, chosen = md.autotune(th_last)
    md.profile(True, every=PROF_EVERY)
    md.timing()  # drains the stream: nothing in flight when the clock starts
    t0 = time.perf_counter()
    out = md.rNNGP(nSamples=steps, seed=12, **start_t)
    tm = md.timing()  # synchronizes the model's stream
    call_dt = time.perf_counter() - t0
    md.close()
    iters_s, dt = loop_rate(tm, steps)
    return {"iters_s": iters_s, "dt": dt, "call_iters_s": steps / call_dt,
            "call_dt": call_dt, "tm": tm, "out": out,
            "beta_mean": out["p.beta.samples"].mean(1),
            "theta_mean": out["p.theta.samples"].mean(1),
            "tune": {"candidates": [list(t) for t in cand],
                     "chosen": list(chosen)}}


def mixed_errors(c, X, y, idx, m, beta):
    """Mixed-precision pass against the oracle's FP64 updateBF and Q at the
    generating theta: max |dB|, max |dF| and the log-likelihood error."""
    import spnngp_amd as nn
    from oracle import oracle as O
    n = len(y)
    theta = [2.0, 0.2, 12.0]
    B0, F0, ld0 = O.updateBF(c.T.reshape(-1).copy(), idx["nnIndx"],
                             idx["nnIndxLU"], n, m, theta, 0,
                             nThreads=cores())
    r = X @ beta - y
    q0 = O.Q(B0, F0, r, r, idx["nnIndx"], idx["nnIndxLU"])
    ll0 = -0.5 * ld0 - 0.5 * q0
    res = {}
    for name, prec in (("mixed", 1), ("fp64", 0)):
        with nn.Model(c, idx["nnIndx"], idx["nnIndxLU"], m, "exponential", X,
                      y, precision=prec) as md:
            o = md.update(theta, slot=0, beta=beta)
            B, F = md.get_bf(0)
        ll = -0.5 * o["logDet"] - 0.5 * o["qResid"]
        res[name] = {"max_abs_dB": float(np.max(np.abs(B - B0))),
                     "max_abs_dF": float(np.max(np.abs(F - F0))),
                     "max_rel_dF": float(np.max(np.abs(F - F0) / F0)),
                     "loglik_rel_err": float(abs(ll - ll0) / abs(ll0))}
    return res


def cores():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(avail, int(env))) if env else avail


def cpu_baseline(c, X, y, idx, m, iters, start):
    """The oracle's loop-for-loop restatement of src/rNNGP.cpp (two updateBF
    per iteration, OpenMP over sites) on this host, for a bounded sample."""
    from oracle import oracle as O
    n, p = X.shape
    nt = cores()
    rng = np.random.default_rng(7)
    tape = O.TapeRng(rng.normal(size=iters * 8 + 16),
                     rng.uniform(size=iters + 4))
    t0 = time.perf_counter()
    O.rNNGP(y, X, p, n, m, c.T.reshape(-1).copy(), 0, idx["nnIndx"],
            idx["nnIndxLU"], rng=tape, nThreads=nt, nSamples=iters, **start)
    dt = time.perf_counter() - t0
    return {"value": iters / dt, "unit": "iter/s", "cores": nt, "cpu_model": cpu_model(),
            "kind": "port",
            "sample": f"{iters} MH iterations of the same n={n}, m={m} "
                      f"chain (oracle rNNGP, OpenMP, {nt} threads), "
                      f"{dt:.1f} s"}


def posterior_samples(S, seed):
    """S stored posterior draws around the truth: beta ~ N(truth, 0.01^2),
    log sigma.sq, log tau.sq and logit phi (on [3, 300]) ~ truth +
    N(0, 0.02^2)."""
    rng = np.random.default_rng(seed)
    beta = rng.normal([1.0, 2.0, -1.0], 0.01, size=(S, 3)).T.copy()
    lphi = np.log((12 - 3) / (300 - 12))
    th = np.vstack([np.exp(np.log(2.0) + rng.normal(0, 0.02, S)),
                    np.exp(np.log(0.2) + rng.normal(0, 0.02, S)),
                    300 - (300 - 3) / (1 + np.exp(lphi
                                                  + rng.normal(0, 0.02, S)))])
    return beta, th


def replicates_bench(c, X, idx, m, S, seed, cpu_samples):
    """rNNGPReplicated (fitted() replicates) for S posterior samples of the
    fitted response model: per sample an updateBF pass and the sequential
    NNGP draw (level-scheduled on the device), plus X beta.  The normals are
    the library's counter-based stream.  The CPU leg is the oracle's
    restatement on `cpu_samples` samples of the same model."""
    import spnngp_amd as nn
    n, p = X.shape
    beta, th = posterior_samples(S, seed)
    args = (X, p, n, m, c, 0, idx["nnIndx"], idx["nnIndxLU"])
    nn.rNNGPReplicated(*args, beta.T[:2].copy(), th[:, :2].copy(), 2,
                       seed=seed)  # warm-up
    t0 = time.perf_counter()
    rep = nn.rNNGPReplicated(*args, beta.T.copy(), th, S,
                             seed=seed)["y.rep.samples"]
    dt = time.perf_counter() - t0
    out = {"n": n, "m": m, "samples": S, "seconds": dt,
           "replicates_per_s": S / dt,
           "site_samples_per_s": n * S / dt,
           "finite": bool(np.all(np.isfinite(rep)))}
    del rep
    if cpu_samples > 0:
        from oracle import oracle as O
        nt = cores()
        z = np.random.default_rng(seed).normal(size=n * cpu_samples)
        t0 = time.perf_counter()
        O.rNNGPReplicated(X.T.reshape(-1).copy(), p, n, m,
                          c.T.reshape(-1).copy(), 0, idx["nnIndx"],
                          idx["nnIndxLU"],
                          # beta is p x S here; the oracle takes S x p
                          # column-major, i.e. this p x S block row-major
                          beta[:, :cpu_samples].copy().reshape(-1),
                          th[:, :cpu_samples].T.reshape(-1).copy(),
                          cpu_samples, O.TapeRng(z, []), nThreads=nt)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_samples / cdt,
                               "unit": "replicates/s", "cores": nt, "cpu_model": cpu_model(),
                               "kind": "port",
                               "sample": f"{cpu_samples} replicate(s) of the "
                                         f"same n={n}, m={m} model (oracle "
                                         f"rNNGPReplicated, OpenMP updateBF, "
                                         f"serial draw), {cdt:.1f} s"}
    return out


def knn_bench(c, idx, m, gpu_s, cpu):
    """mkNNIndxCB at the bench size: the device search's wall time (host
    strip sort and copies included; a second, warm call) and, with cpu, the oracle's restatement of the
    reference's code-book search (src/nn.cpp:128-257, OpenMP over sites) on
    the same sites, whose neighbor lists must equal the device's bit for
    bit."""
    import spnngp_amd as nn
    t0 = time.perf_counter()
    nn.mkNNIndxCB(c, m)  # warm: the first call also loads the module
    out = {"n": int(c.shape[0]), "m": m, "gpu_s": time.perf_counter() - t0,
           "gpu_first_call_s": gpu_s}
    if cpu:
        from oracle import oracle as O
        nt = cores()
        t0 = time.perf_counter()
        ref = O.mkNNIndxCB(c, m, nt)
        cdt = time.perf_counter() - t0
        out["bit_exact"] = bool(
            np.array_equal(ref["nnIndx"], idx["nnIndx"])
            and np.array_equal(ref["nnIndxLU"], idx["nnIndxLU"])
            and np.array_equal(ref["nnDist"], idx["nnDist"]))
        out["cpu_baseline"] = {
            "value": cdt, "unit": "s", "cores": nt, "cpu_model": cpu_model(), "kind": "port",
            "sample": f"the whole n={c.shape[0]}, m={m} search (oracle "
                      f"mkNNIndxCB, OpenMP)"}
    return out


def latent_bench(c, X, y, idx, m, iters, site_iters, cpu_iters, seed):
    """Latent model (sNNGP) on the same data: the neighbors-of-neighbors
    index (mkUIndx, search type 2) and MCMC iterations/s with the coloured
    sweep (colour classes in parallel) and with the site-order sweep (the
    reference's scan, level by level).  The sweep's normals are the
    device's counter-based stream.  The CPU leg is the oracle's sNNGP
    (OpenMP updateBF, serial sweep) on `cpu_iters` iterations, fed by the
    library's R generator through C callbacks."""
    import spnngp_amd as nn
    n, p = X.shape
    t0 = time.perf_counter()
    U = nn.mkUIndx(n, m, idx["nnIndx"], idx["nnIndxLU"], 2)
    u_s = time.perf_counter() - t0
    beta0 = np.linalg.lstsq(X, y, rcond=None)[0]
    args = (y, X, p, n, m, c, 0, idx["nnIndx"], idx["nnIndxLU"], U["u.indx"],
            U["u.indx.lu"], U["ui.indx"], [2.0, 2.0], [2.0, 0.2],
            [3.0, 300.0], None, beta0, 2.0, 0.2, 12.0, 0.0, 0.0, 0.0, 0.02,
            0.0)
    out = {"n": n, "m": m, "mkUIndx_s": u_s}
    for sweep, k in (("coloured", iters), ("site", site_iters)):
        if k <= 0:
            continue
        nn.sNNGP(*args, 1, seed=seed, sweep=sweep)  # schedule + first pass
        t0 = time.perf_counter()
        nn.sNNGP(*args, 1, seed=seed, sweep=sweep)
        t1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = nn.sNNGP(*args, k + 1, seed=seed, sweep=sweep)
        tk = time.perf_counter() - t0
        per = (tk - t1) / k  # set-up (schedule, uploads, first pass) removed
        out[sweep] = {"iters_s": 1.0 / per, "ms_per_iter": per * 1e3,
                      "setup_s": t1 - per, "iters": k,
                      "acceptance": r["accept"] / (k + 1),
                      "finite": bool(np.all(np.isfinite(r["p.w.samples"])))}
        del r
    if cpu_iters > 0:
        from oracle import oracle as O
        nt = cores()
        t0 = time.perf_counter()
        Uo = O.mkUIndx(n, m, idx["nnIndx"], idx["nnIndxLU"], 2)
        cu = time.perf_counter() - t0
        assert all(np.array_equal(Uo[k], U[k]) for k in U)
        rr = nn.RRng(seed)
        t0 = time.perf_counter()
        O.sNNGP(y, X, p, n, m, c.T.reshape(-1).copy(), 0, idx["nnIndx"],
                idx["nnIndxLU"], U["u.indx"], U["u.indx.lu"], U["ui.indx"],
                [2.0, 2.0], [2.0, 0.2], [3.0, 300.0], None, beta0, 2.0, 0.2,
                12.0, 0.0, 0.02, 0.0, cpu_iters, rr, nThreads=nt)
        cdt = time.perf_counter() - t0
        rr.close()
        out["cpu_baseline"] = {
            "value": cpu_iters / cdt, "unit": "iter/s", "cores": nt, "cpu_model": cpu_model(),
            "kind": "port", "mkUIndx_s": cu,
            "sample": f"{cpu_iters} iteration(s) of the same n={n}, m={m} "
                      f"latent chain (oracle sNNGP: OpenMP updateBF, serial "
                      f"sweep), {cdt:.1f} s; mkUIndx search type 2 on one "
                      f"core, {cu:.2f} s"}
    return out


def kriging_serving(c, X, y, m, seconds, rate, S, seed, cpu_sites=0):
    """Kriging server under load: requests of q ~ floor(exp(U[0, ln 1e5]))
    new sites, Poisson arrivals at `rate` per second for `seconds`, against
    the resident fitted model with S posterior samples.  Latency is result
    on the host minus the scheduled arrival time (so a late client counts
    against the server)."""
    import threading
    import spnngp_amd as nn
    rng = np.random.default_rng(seed)
    n, p = X.shape
    beta, th = posterior_samples(S, seed)
    POOL = 100_000
    pool_c = rng.uniform(size=(2, POOL))
    pool_X = np.vstack([np.ones(POOL), rng.normal(size=(p - 1, POOL))])
    t0 = time.perf_counter()
    srv = nn.Server(X, y, c, m, beta, th, "exponential",
                    device_bytes=16 << 30, host_bytes=6 << 30,
                    batch_sites=16384, depth=3)
    setup_s = time.perf_counter() - t0

    def req(q, off):
        return (np.ascontiguousarray(pool_X[:, off:off + q].T),
                np.ascontiguousarray(pool_c[:, off:off + q].T))

    # unloaded latency of a single 10k-site request (same delivery as under
    # load: result in pinned host memory)
    alone = []
    for r in range(11):
        X0, c0 = req(10_000, 37 * r)
        t = srv.submit(X0, c0, 10_000_000 + r)
        _, info = srv.wait(t, timeout=60.0, copy=False)
        srv.release(t)
        if r:
            alone.append((info["t_done"] - info["t_submit"]) * 1e3)

    nreq = rng.poisson(rate * seconds)
    arrivals = np.sort(rng.uniform(0, seconds, nreq))
    sizes = np.floor(np.exp(rng.uniform(0, np.log(1e5), nreq))).astype(int)
    sizes = np.clip(sizes, 1, POOL)
    offs = rng.integers(0, POOL - sizes + 1)
    tickets = [None] * nreq
    done = {}
    late = [0.0]
    base = time.monotonic() + 0.2

    def producer():
        for k in range(nreq):
            X0, c0 = req(int(sizes[k]), int(offs[k]))
            dt = base + arrivals[k] - time.monotonic()
            if dt > 0:
                time.sleep(dt)
            else:
                late[0] = max(late[0], -dt)
            tickets[k] = srv.submit(X0, c0, k)

    def consumer():
        for _ in range(nreq):
            t, info = srv.wait(-1, timeout=120.0, copy=False)
            done[t] = info
            srv.release(t)

    th_p = threading.Thread(target=producer)
    th_c = threading.Thread(target=consumer)
    wall0 = time.monotonic()
    th_p.start()
    th_c.start()
    th_p.join()
    th_c.join()
    wall = time.monotonic() - wall0
    st = srv.stats()
    srv.close()
    lat = np.empty(nreq)
    ok = 0
    for k in range(nreq):
        info = done[tickets[k]]
        ok += info["status"] == 0
        lat[k] = (info["t_done"] - (base + arrivals[k])) * 1e3
    mid = (sizes >= 5_000) & (sizes <= 20_000)
    # the kriging kernel's own rate: device time of the retired launches
    # (HIP events around each launch on the server's kernel stream)
    kflops = st["kernel_site_samples"] * f_krig(m, p)
    kachieved = kflops / (st["kernel_ms"] * 1e-3) / 1e12 \
        if st["kernel_ms"] > 0 else 0.0
    roof = roofline_obj(
        "kriging (per-lane kernel, m=15, exponential, batched server "
        "launches of <= 16384 sites x %d samples)" % S, kachieved,
        *traffic("kriging_server_launch_n1e6_m15_S500",
                 "nngp::k_pred_lane<15, 0>"),
        flops_per_site_sample=f_krig(m, p),
        kernel_ms_total=st["kernel_ms"], launches=int(st["launches"]),
        mean_launch_ms=st["kernel_ms"] / max(st["launches"], 1))
    cpu = None
    if cpu_sites > 0:
        # the oracle's rNNGPPredict (src/rNNGPPredict.cpp:25-222, OpenMP over
        # samples as the reference) on cpu_sites new sites x S samples
        from oracle import oracle as O
        X0, c0 = req(cpu_sites, 0)
        nn0, _ = nn.nn2(c, c0, m)
        z = rng.normal(size=cpu_sites * S)
        nt = cores()
        t0 = time.perf_counter()
        O.rNNGPPredict(X.T.reshape(-1).copy(), y, c.T.reshape(-1).copy(), n,
                       p, m, X0.T.reshape(-1).copy(), c0.T.reshape(-1).copy(),
                       cpu_sites, nn0.T.reshape(-1).copy(),
                       beta.T.reshape(-1).copy(), th.T.reshape(-1).copy(), S,
                       0, z, nThreads=nt)
        cdt = time.perf_counter() - t0
        cpu = {"value": cpu_sites * S / cdt, "unit": "site_samples/s",
               "cores": nt, "cpu_model": cpu_model(), "kind": "port",
               "q1e4_request_ms": cdt * 1e3 * 10_000 / cpu_sites,
               "sample": f"{cpu_sites} new sites x {S} samples of the same "
                         f"model (oracle rNNGPPredict, OpenMP), {cdt:.2f} s; "
                         f"the 10k-site figure scales it linearly"}
    return {"requests": int(nreq), "ok": int(ok), "rate_per_s": rate,
            "seconds": seconds, "samples": S,
            "q": "floor(exp(U[0, ln 1e5])), mean %.0f" % sizes.mean(),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "p50_ms_q5e3_2e4": float(np.percentile(lat[mid], 50)),
            "p99_ms_q5e3_2e4": float(np.percentile(lat[mid], 99)),
            "site_samples_per_s": float(st["site_samples"] / wall),
            "launches": int(st["launches"]),
            "device_bytes_peak": int(st["device_bytes_peak"]),
            "client_late_ms_max": late[0] * 1e3,
            "alone_q1e4_p50_ms": float(np.percentile(alone, 50)),
            "alone_q1e4_p99_ms": float(np.percentile(alone, 99)),
            "setup_s": setup_s,
            "roofline": roof,
            "cpu_baseline": cpu,
            "note": "resident model (n=1e6, m=15, exponential), nn2 + "
                    "kriging on the GPU, results delivered to pinned host "
                    "memory; latency = result on host - scheduled arrival"}


def conj_grid_bench(n, m, k, seed, cpu_points):
    """spConjNNGP's k-fold grid search at scale: n sites ~ U[0,1]^2, X = 3
    N(0,1) columns (no intercept), beta = (1, 2, -1), w from the NNGP prior
    (matern nu = 1.5, sigma.sq = 1, phi = 8), tau.sq = 0.1; grid phi in 25
    log-spaced points of [1, 40] x alpha in 20 log-spaced points of
    [0.01, 1] (nu = 1.5); k folds of floor(n/k) held-out sites each."""
    import spnngp_amd as nn
    rng = np.random.default_rng(seed)
    c = rng.uniform(size=(n, 2))
    c = c[np.argsort(c[:, 0], kind="stable")]
    X = rng.normal(size=(n, 3))
    idx = nn.mkNNIndxCB(c, m)
    B, F, _ = nn.updateBF(c, idx["nnIndx"], idx["nnIndxLU"], m,
                          [1.0, 0.0, 8.0, 1.5], "matern")
    w = nn.prior_draw(B, F, idx["nnIndx"], idx["nnIndxLU"],
                      rng.normal(size=n))
    y = X @ [1.0, 2.0, -1.0] + w + rng.normal(scale=np.sqrt(0.1), size=n)
    nho = n // k
    ho = np.array([np.sort(rng.choice(n, nho, replace=False))
                   for _ in range(k)], dtype=np.int32)
    phis = np.exp(np.linspace(np.log(1.0), np.log(40.0), 25))
    alphas = np.exp(np.linspace(np.log(0.01), np.log(1.0), 20))
    ta = np.array([[ph, al, 1.5] for ph in phis for al in alphas])
    g = ta.shape[0]
    t0 = time.perf_counter()
    cv = nn.CvGrid(y, X, c, m, ho, [2.0, 1.0], "matern")
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    first = cv.run(ta)  # autotune + graph capture + run
    first_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = cv.run(ta)
    wall_s = time.perf_counter() - t0
    det = cv.run(ta, details=True) if cpu_points > 0 else None  # untimed
    cv.close()
    best = int(np.argmin(r["scores"][:, 1]))
    nmod = n - nho
    # flops: updateConjBF + cross-products per training site, kriging-like
    # solve + p-vector variance per held-out site (exp and sqrt count 1)
    f_site = f_alg(m) + 2 * m * 4 + 10
    f_pred = f_alg(m) + 2 * 3 * m + 3 * 3 + 20
    flops = k * g * (nmod * f_site + nho * f_pred)
    res = {"n": n, "m": m, "k_fold": k, "grid": g, "problems": k * g,
           "site_updates": k * g * nmod, "predictions": k * g * nho,
           "graph_ms": r["graph_ms"], "run_wall_s": wall_s,
           "first_run_s": first_s, "setup_s": setup_s,
           "site_updates_per_s": k * g * nmod / (r["graph_ms"] * 1e-3),
           "achieved_tflops": flops / (r["graph_ms"] * 1e-3) / 1e12,
           "bf_config": list(r["bf_config"]),
           "tune_ms": [[t[0], t[1], t[2]] for t in first["tune"]],
           "best_point": {"phi": float(ta[best, 0]),
                          "alpha": float(ta[best, 1]),
                          "crps": float(r["scores"][best, 1])}}
    if cpu_points > 0:
        from oracle import oracle as O
        keep = np.ones(n, bool)
        keep[ho[0]] = False
        cm, Xm, ym = c[keep], X[keep], y[keep]
        c0, X0 = c[ho[0]], X[ho[0]]
        nnm = nn.mkNNIndxCB(cm, m)
        nn0, _ = nn.nn2(cm, c0, m)
        sub = ta[:cpu_points]
        nt = cores()
        t0 = time.perf_counter()
        ref = O.cNNGP(ym, Xm.T.reshape(-1).copy(), 3, nmod,
                cm.T.reshape(-1).copy(), sub.reshape(-1).copy(),
                nnm["nnIndx"], nnm["nnIndxLU"], X0.T.reshape(-1).copy(),
                c0.T.reshape(-1).copy(), nho, nn0.T.reshape(-1).copy(),
                cpu_points, m, [2.0, 1.0], 2, nThreads=nt)
        dt = time.perf_counter() - t0
        # the device's fold 1 at the same grid points, against the oracle
        bh = np.asarray(ref["beta.hat"]).T  # g x p
        abo = np.asarray(ref["ab"]).T       # g x 2
        res["cpu_check"] = {
            "problems": cpu_points,
            "beta_hat_max_rel": float(np.max(
                np.abs(det["beta.hat"][0, :cpu_points] - bh) / np.abs(bh))),
            "ab_max_rel": float(np.max(
                np.abs(det["ab"][0, :cpu_points] - abo) / np.abs(abo)))}
        res["cpu_baseline"] = {
            "value": cpu_points * nmod / dt, "unit": "site-updates/s",
            "cores": nt, "cpu_model": cpu_model(), "kind": "port",
            "sample": f"oracle cNNGP, fold 1 x first {cpu_points} grid "
                      f"points ({nmod} training + {nho} held-out sites), "
                      f"{dt:.1f} s; the full grid would take "
                      f"{dt * k * g / cpu_points:.0f} s"}
    return res


def stress_bench(n, m, iters, seed):
    """The stress case: n sites, 90% from 50 Gaussian clusters (centres
    U[0,1]^2, sd 0.02) and 10% U[0,1]^2, then x stretched by 3; X = [1,
    N(0,1)], beta = (1, 5); w from the NNGP prior (matern nu = 2.5, phi =
    20, sigma.sq = 1, plus a 1e-6 jitter), tau.sq = 0.05; m neighbors.  nu
    is sampled, so every pass evaluates the general Bessel-K matern (one
    site per wavefront).  Times `iters` MH iterations after one warm-up
    iteration, in FP64 and in the mixed form (each autotuned), and reports
    the mixed pass's error against the FP64 pass at the generating theta."""
    import spnngp_amd as nn
    rng = np.random.default_rng(seed)
    nc = int(0.9 * n)
    centres = rng.uniform(size=(50, 2))
    c = np.vstack([centres[rng.integers(0, 50, nc)]
                   + rng.normal(scale=0.02, size=(nc, 2)),
                   rng.uniform(size=(n - nc, 2))])
    c[:, 0] *= 3.0
    c = c[np.argsort(c[:, 0], kind="stable")]
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    t0 = time.perf_counter()
    idx = nn.mkNNIndxCB(c, m)
    knn_s = time.perf_counter() - t0
    # the draw carries a 1e-6 nugget: with sites ~1e-4 apart inside a
    # cluster, the nugget-free nu = 2.5 neighbor covariances are singular to
    # working precision (dpotrf fails on them, as the reference's would)
    B, F, _ = nn.updateBF(c, idx["nnIndx"], idx["nnIndxLU"], m,
                          [1.0, 1e-6, 20.0, 2.5], "matern")
    w = nn.prior_draw(B, F, idx["nnIndx"], idx["nnIndxLU"],
                      rng.normal(size=n))
    del B, F
    y = X @ [1.0, 5.0] + w + rng.normal(scale=np.sqrt(0.05), size=n)
    start = dict(sigmaSqIG=[2.0, 1.0], tauSqIG=[2.0, 0.05],
                 phiUnif=[3.0, 100.0], nuUnif=[0.5, 4.0],
                 betaStarting=np.linalg.lstsq(X, y, rcond=None)[0],
                 sigmaSqStarting=1.0, tauSqStarting=0.05, phiStarting=20.0,
                 nuStarting=2.5, sigmaSqTuning=0.01, tauSqTuning=0.01,
                 phiTuning=0.01, nuTuning=0.01)
    th0 = [1.0, 0.05, 20.0, 2.5]
    beta0 = start["betaStarting"]
    legs = {}
    ref = None
    for prec, name in ((0, "fp64"), (1, "mixed")):
        md = nn.Model(c, idx["nnIndx"], idx["nnIndxLU"], m, "matern", X, y,
                      precision=prec)
        md.rNNGP(nSamples=1, seed=21, **start)
        cand, chosen = md.autotune([1.0, 0.05, 20.0, 2.4])
        # one pass at the generating theta: the mixed form's error against
        # the FP64 pass on the same inputs
        po = md.update(th0, slot=0, beta=beta0)
        ll = -0.5 * po["logDet"] - 0.5 * po["qResid"]
        B, F = md.get_bf(0)
        md.profile(True)
        md.timing()
        t0 = time.perf_counter()
        out = md.rNNGP(nSamples=iters, seed=22, **start)
        tm = md.timing()
        call_dt = time.perf_counter() - t0
        md.close()
        bf_ms = tm["bf_ms"] / max(tm["bf_launches"], 1)
        pairs = m * (m + 1) // 2
        it_s, dt = loop_rate(tm, iters)
        leg = {"iters_s": it_s, "ms_per_iter": dt / iters * 1e3,
               "call_iters_s": iters / call_dt,
               "updateBF_ms": bf_ms,
               "bessel_evals_per_s": n * pairs / (bf_ms * 1e-3),
               "flops_per_s": n * f_alg(m) / (bf_ms * 1e-3),
               "bf_tune": {"candidates": [list(t) for t in cand],
                           "chosen": list(chosen)},
               "accept": out["accept"]}
        if ref is None:
            ref = (ll, B, F)
        else:
            leg["vs_fp64"] = {
                "loglik_rel_err": abs(ll - ref[0]) / abs(ref[0]),
                "max_rel_dF": float(np.max(np.abs(F - ref[2]) / ref[2])),
                "max_dB_over_max_B": float(np.max(np.abs(B - ref[1]))
                                           / np.max(np.abs(ref[1])))}
        del B, F
        legs[name] = leg
    del ref
    fp, mx = legs["fp64"], legs["mixed"]
    fam, th = fp["bf_tune"]["chosen"]
    kname = (f"nngp::k_bf_grp<32, true> ({th} threads)" if fam == "row" else
             f"nngp::k_bf_coop<false, 32, false> ({th} threads)")
    fp["roofline"] = roofline_obj(
        "updateBF (%s kernel, m=32, Bessel-K matern, %d threads)"
        % ("16-lane-row" if fam == "row" else "wavefront", th),
        fp["flops_per_s"] / 1e12,
        *(traffic("updateBF_stress_n4e6_m32", kname)
          if (n, m) == (4_000_000, 32) else (None, "other workload")),
        algorithmic_bytes=(24 + 4 * m) * n,
        flops_per_launch=n * f_alg(m), launch_ms=fp["updateBF_ms"])
    return {"n": n, "m": m, "iters": iters, "knn_s": knn_s,
            # FP64 leg at the top level, the mixed leg beside it
            **fp, "mixed": mx,
            "mixed_speedup": mx["iters_s"] / fp["iters_s"],
            "precision": "fp64 and mixed (FP32 build incl. Bessel K, FP32 "
                         "factor and solves, one FP64 refinement), each "
                         "autotuned"}


def roofline_obj(kernel, achieved_tflops, traffic_bytes, traffic_note,
                 **extra):
    """The roofline record of an FP64 compute-bound kernel: the kernels run
    on the FP64 vector ALUs (no matrix cores), whose peak is 78.6 TFLOP/s."""
    return {"bound": "fp64_valu",
            "bound_note": "FP64 vector FMA issue (no MFMA); peak 78.6 TFLOP/s",
            "kernel": kernel, "achieved": achieved_tflops,
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP64_TFLOPS,
            "traffic": traffic_bytes, "traffic_note": traffic_note, **extra}


def chain_kernel(chosen, m, mixed=False, zs=4):
    """Name of the chain's updateBF kernel (exponential) for an autotuned
    shape of the FP64 or the mixed form."""
    fam, threads = chosen
    M = 15 if m == 15 else m
    if mixed:
        if fam == "lane+reg-copy":
            # rows staged (zs / 2 chunks) when the block's stage fits 64 KiB
            stage = threads // 64 * (M + 1) * (zs // 2) * 64 * 16
            zh = zs // 2 if m > 8 and stage <= 65536 else 0
            return f"nngp::k_bf_mixed_reg<{M}, 0, {zh}> ({threads} threads)"
        if fam in ("lane", "lane+lds-copy"):
            copy = "true" if fam == "lane+lds-copy" else "false"
            return f"nngp::k_bf_mixed<{M}, 0, {copy}> ({threads} threads)"
        return f"{fam} ({threads} threads)"
    if fam == "lane+lds-z":
        return f"nngp::k_bf_lane_lds<{M}, 0, 2, false> ({threads} threads)"
    if fam == "lane":
        return f"nngp::k_bf_lane<{M}, 0, false> ({threads} threads)"
    return f"{fam} ({threads} threads)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--m", type=int, default=15)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-iters", type=int, default=40)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--grid-n", type=int, default=500_000)
    ap.add_argument("--grid-cpu-points", type=int, default=4)
    ap.add_argument("--stress-n", type=int, default=4_000_000)
    ap.add_argument("--stress-iters", type=int, default=5)
    ap.add_argument("--rep-samples", type=int, default=500)
    ap.add_argument("--rep-cpu-samples", type=int, default=1)
    ap.add_argument("--latent-iters", type=int, default=100)
    ap.add_argument("--latent-site-iters", type=int, default=3)
    ap.add_argument("--latent-cpu-iters", type=int, default=2)
    ap.add_argument("--knn-cpu", type=int, default=1)
    ap.add_argument("--serve-cpu-sites", type=int, default=5000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is "
              "not supported (the multi-GPU sharding is a design only)",
              file=sys.stderr)
        sys.exit(2)

    import spnngp_amd as nn
    if nn.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    n, m = args.n, args.m
    print(f"[bench] generating n={n} m={m}", file=sys.stderr, flush=True)
    c, X, y, idx, knn_s = make_data(n, m, args.seed)
    p = X.shape[1]
    beta0 = np.linalg.lstsq(X, y, rcond=None)[0]
    start = dict(sigmaSqIG=[2.0, 2.0], tauSqIG=[2.0, 0.2], phiUnif=[3.0, 300.0],
                 nuUnif=None, betaStarting=beta0, sigmaSqStarting=2.0,
                 tauSqStarting=0.2, phiStarting=12.0, nuStarting=0.0,
                 sigmaSqTuning=0.02, tauSqTuning=0.02, phiTuning=0.02,
                 nuTuning=0.0)
    print(f"[bench] FP64 chain, warmup {args.warmup}", file=sys.stderr,
          flush=True)
    f64 = mcmc_bench(c, X, y, idx, m, args.steps, args.warmup, start, 0)
    print("[bench] mixed chain", file=sys.stderr, flush=True)
    mix = mcmc_bench(c, X, y, idx, m, args.steps, args.warmup, start, 1)
    mix_err = mixed_errors(c, X, y, idx, m, f64["beta_mean"])
    iters_s, dt, tm, out = f64["iters_s"], f64["dt"], f64["tm"], f64["out"]
    bf_ms = tm["bf_ms"] / max(tm["bf_launches"], 1)
    flops = sum(f_alg(min(i, m)) for i in range(1, min(n, m))) \
        + (n - m) * f_alg(m)
    achieved = flops / (bf_ms * 1e-3) / 1e12
    mix_ms = mix["tm"]["bf_ms"] / max(mix["tm"]["bf_launches"], 1)
    mix_achieved = flops / (mix_ms * 1e-3) / 1e12
    mixed = {"iters_s": mix["iters_s"], "call_iters_s": mix["call_iters_s"],
             "speedup_vs_fp64": mix["iters_s"] / iters_s,
             "roofline": dict(roofline_obj(
                 "updateBF (mixed per-lane kernel, m=15, exponential, %s %d "
                 "threads)" % tuple(mix["tune"]["chosen"]), mix_achieved,
                 *(traffic("updateBF_chain_mixed_n1e6_m15",
                           chain_kernel(mix["tune"]["chosen"], m, True,
                                        p + 1))
                   if (n, m) == (1_000_000, 15) else
                   (None, "other workload")),
                 algorithmic_bytes=(24 + 4 * m) * n,
                 algorithmic_bytes_fused=(24 + 4 * m + 8 * (p + 1)) * n,
                 flops_per_launch=flops,
                 launch_ms=mix_ms, launches=mix["tm"]["bf_launches"]),
                 bound="valu",
                 bound_note="FP32 build and factor, FP64 refinement and "
                            "sums; non-packed FP32 and FP64 vector FMAs "
                            "issue at the same rate, 78.6 TFLOP/s of the "
                            "algorithm's flops"),
             "updateBF_ms": mix["tm"]["bf_ms"] / max(mix["tm"]["bf_launches"],
                                                      1),
             "acceptance": mix["out"]["accept"] / args.steps,
             "posterior_mean_diff": {
                 "beta": (mix["beta_mean"] - f64["beta_mean"]).tolist(),
                 "theta": (mix["theta_mean"] - f64["theta_mean"]).tolist()},
             "errors_vs_oracle": mix_err,
             "bf_tune": mix["tune"]}

    print("[bench] cpu baseline", file=sys.stderr, flush=True)
    cpu = cpu_baseline(c, X, y, idx, m, args.cpu_iters, start) \
        if args.cpu_iters > 0 else None
    print("[bench] kriging", file=sys.stderr, flush=True)
    krig = kriging_serving(c, X, y, m, args.serve_seconds, args.serve_rate,
                           500, 4, args.serve_cpu_sites) \
        if args.serve_seconds > 0 else None

    print("[bench] replicates", file=sys.stderr, flush=True)
    reps = replicates_bench(c, X, idx, m, args.rep_samples, 6,
                            args.rep_cpu_samples) \
        if args.rep_samples > 0 else None

    print("[bench] neighbor search", file=sys.stderr, flush=True)
    knn = knn_bench(c, idx, m, knn_s, args.knn_cpu > 0)

    print("[bench] latent", file=sys.stderr, flush=True)
    latent = latent_bench(c, X, y, idx, m, args.latent_iters,
                          args.latent_site_iters, args.latent_cpu_iters, 7) \
        if args.latent_iters > 0 else None

    print("[bench] conjugate grid", file=sys.stderr, flush=True)
    grid = conj_grid_bench(args.grid_n, 15, 5, 3, args.grid_cpu_points) \
        if args.grid_n > 0 else None

    print("[bench] stress", file=sys.stderr, flush=True)
    stress = stress_bench(args.stress_n, 32, args.stress_iters, 5) \
        if args.stress_n > 0 else None

    line = {
        "metric": METRIC, "value": iters_s, "unit": "iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded U[0,1]^2 coords, Gaussian X, w from the "
                "NNGP prior",
        "config": {"workload": "response NNGP MCMC (rNNGP), n=1e6, m=15, "
                               "exponential, p=3, FP64",
                   "n": n, "m": m, "p": p, "cov_model": "exponential",
                   "precision": "fp64", "tuning": 0.02},
        "acceptance": out["accept"] / args.steps,
        "roofline": roofline_obj(
            "updateBF (per-lane kernel, m=15, exponential, %s %d threads)"
            % tuple(f64["tune"]["chosen"]), achieved,
            *(traffic("updateBF_chain_n1e6_m15",
                      chain_kernel(f64["tune"]["chosen"], m))
              if (n, m) == (1_000_000, 15) else (None, "other workload")),
            algorithmic_bytes=(24 + 4 * m) * n,
            # the fused pass also reads each site's own [X | y] row for the
            # cross-products (neighbors' rows are gathers of the same data)
            algorithmic_bytes_fused=(24 + 4 * m + 8 * (p + 1)) * n,
            flops_per_launch=flops,
            launch_ms=bf_ms, launches=tm["bf_launches"]),
        # one updateBF and one sum launch per step; the call's initial pass
        # adds one of each outside the loop, so per-launch means
        "breakdown_ms_per_step": {
            "updateBF": bf_ms,
            "sums": tm["sum_ms"] / max(tm["sum_launches"], 1),
            "host_and_sync": dt * 1e3 / args.steps - bf_ms
                             - tm["sum_ms"] / max(tm["sum_launches"], 1)},
        "call_iters_s": f64["call_iters_s"],
        "cpu_baseline": cpu,
        "kriging": krig,
        "conj_grid": grid,
        "stress": stress,
        "replicates": reps,
        "latent": latent,
        "mixed": mixed,
        "bf_tune": f64["tune"],
        "knn_s": knn_s,
        "knn": knn,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
