"""R-level front ends of spNNGP's hot path: spNNGP(method="response" or
"latent") (R/spNNGP.R), spConjNNGP() (R/spConjNNGP.R), predict() and fitted() for
both (R/generics.R:583-675, :1-268) and the CV scores crps() / rmspe()
(R/util.R:247-255).

These mirror the R functions' argument checks, defaults and host-side
bookkeeping (x-ordering, OLS starting beta, holdout folds, argmin of the
k-fold scores, sub-sampling, putting results back in the caller's order);
every O(n) step runs in libnngp_hip.so.  Argument names follow R with dots
turned into underscores; lists become dicts with R's names.

Differences from R that a caller sees:
  * R itself is not called: holdout folds, the exact conjugate samples and
    fitted()'s parameter draws use numpy's Generator (seeded by `seed`),
    replicate normals the library's counter-based stream, and the sampler
    the library's built-in generator unless `rng` passes a host random
    source (a nngp_rng-compatible object).  rng=RRng(s), the library's
    copy of R's default generator, gives the stream of R's set.seed(s).
    Given the same draws the results are the reference's.
  * method = "response" and "latent" with family = "gaussian" are on this
    path, with predict() and fitted() for both (the binomial and SLGP
    models are not).
"""
from __future__ import annotations

import time

import numpy as np
from scipy import special as _sp

from . import nngp as _nn

COV_MODELS = _nn.COV_MODELS


# ---------------------------------------------------------------------------
# scores (R/util.R:247-255)
# ---------------------------------------------------------------------------
def crps(y, y_hat, y_var):
    """Mean continuous ranked probability score of Gaussian predictions,
    oriented as R/util.R:247-251 (lower is better)."""
    y, y_hat, y_var = (np.asarray(a, dtype=np.float64) for a in
                       (y, y_hat, y_var))
    sd = np.sqrt(y_var)
    z = (y - y_hat) / sd
    dn = np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)
    pn = _sp.ndtr(z)
    return float(np.mean(-sd * (1 / np.sqrt(np.pi) - 2 * dn
                                - z * (2 * pn - 1))))


def rmspe(y, y_hat):
    """Root mean squared prediction error (R/util.R:253-255)."""
    y, y_hat = np.asarray(y, float), np.asarray(y_hat, float)
    return float(np.sqrt(np.mean((y - y_hat) ** 2)))


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _order(coords, ord):
    n = coords.shape[0]
    if ord is None:
        return np.argsort(coords[:, 0], kind="stable")  # order(coords[,1])
    ord = np.asarray(ord, dtype=np.int64)
    if ord.shape != (n,):
        raise ValueError("error: supplied order vector ord must be of "
                         "length n")
    return ord


def _lower(d):
    return {str(k).lower(): v for k, v in (d or {}).items()}


def _pair(v, name, positive=True, increasing=False):
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    if v.size != 2:
        raise ValueError(f"error: {name} must be a vector of length 2")
    if positive and np.any(v <= 0):
        raise ValueError(f"error: {name} must be a positive vector of "
                         "length 2" + (" with element 1 < element 2"
                                       if increasing else ""))
    if increasing and not v[0] < v[1]:
        raise ValueError(f"error: {name} must be a positive vector of "
                         "length 2 with element 1 < element 2")
    return v


def _cov_name(cov_model):
    if cov_model not in COV_MODELS:
        raise ValueError(f"error: specified cov.model '{cov_model}' is not a "
                         "valid option; choose from "
                         + ", ".join(COV_MODELS) + ".")
    return cov_model


def _neighbors(coords, m, search_type):
    if search_type not in ("brute", "cb"):
        raise ValueError(f"error: specified search.type '{search_type}' is "
                         "not a valid option; choose from brute, cb.")
    t0 = time.perf_counter()
    idx = (_nn.mkNNIndx if search_type == "brute" else _nn.mkNNIndxCB)(
        coords, m)
    idx["run.time"] = time.perf_counter() - t0
    return idx


def n_indx_list(nn_indx, nn_indx_lu, n):
    """mk.n.indx.list of R/util.R:272-279: 1-based neighbor sets per site
    (None for the first)."""
    out = [None]
    for i in range(1, n):
        off, cnt = nn_indx_lu[i], nn_indx_lu[n + i]
        out.append(np.asarray(nn_indx[off:off + cnt]) + 1)
    return out


def _sub_sample(sub_sample, n_samples, default_start):
    ss = dict(sub_sample or {})
    start = ss.get("start", default_start)
    end = ss.get("end", n_samples)
    thin = ss.get("thin", 1)
    if start >= n_samples:
        raise ValueError("error: in sub.sample, invalid start")
    if end > n_samples:
        raise ValueError("error: in sub.sample, invalid end")
    if thin >= n_samples:
        raise ValueError("error: in sub.sample, invalid thin")
    s_indx = np.arange(int(start), int(end) + 1, int(thin))  # 1-based
    return {"start": start, "end": end, "thin": thin}, s_indx


# ---------------------------------------------------------------------------
# spNNGP(method = "response")
# ---------------------------------------------------------------------------
def spNNGP(y, X, coords, *, starting, tuning, priors, n_samples,
           method="response", n_neighbors=15, cov_model="exponential",
           n_omp_threads=1, search_type="cb", ord=None,
           return_neighbor_info=False, neighbor_info=None, fit_rep=False,
           sub_sample=None, verbose=False, n_report=100, rng=None, seed=1,
           precision=0, x_names=None, u_search_type=2, sweep="site"):
    """spNNGP(y ~ X - 1, coords, method = "response" | "latent", ...) of
    R/spNNGP.R.  Returns a dict with R's element names (p.beta.samples
    n.samples x p, p.theta.samples n.samples x nTheta, and for the latent
    model p.w.samples n x n.samples).  The latent model takes
    u_search_type (R's u.search.type, default 2) and sweep, the order of
    its w sweep ("site": the reference's; "coloured": colour classes of the
    moral graph, a parallel Gibbs scan)."""
    method = method.lower()
    if method not in ("response", "latent"):
        raise ValueError("error: specified method '" + method
                         + "' is not a valid option; choose from response, "
                         "latent")
    latent = method == "latent"
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    coords = np.asarray(coords, dtype=np.float64)
    n, p = X.shape
    if y.shape[0] != n or coords.shape != (n, 2):
        raise ValueError("error: either coords has more than two columns or "
                         "the number of rows is different than data used "
                         "in the model formula")
    cov_model = _cov_name(cov_model)
    matern = cov_model == "matern"

    if neighbor_info is not None:
        need = ("n.neighbors", "nn.indx", "nn.indx.lu", "ord")
        if not all(k in neighbor_info for k in need):
            raise ValueError("The supplied neighbor.info is malformed.")
        ord = np.asarray(neighbor_info["ord"], dtype=np.int64)
        n_neighbors = int(neighbor_info["n.neighbors"])
    ord = _order(coords, ord)
    coords_o, X_o, y_o = coords[ord], X[ord], y[ord]

    pr = _lower(priors)
    if "sigma.sq.ig" not in pr:
        raise ValueError("error: sigma.sq.IG must be specified")
    sigma_sq_IG = _pair(pr["sigma.sq.ig"], "sigma.sq.IG")
    if "tau.sq.ig" not in pr:
        raise ValueError("error: tau.sq.IG must be specified")
    tau_sq_IG = _pair(pr["tau.sq.ig"], "tau.sq.IG")
    if "phi.unif" not in pr:
        raise ValueError("error: phi.Unif must be specified")
    phi_Unif = _pair(pr["phi.unif"], "phi.Unif", increasing=True)
    nu_Unif = None
    if matern:
        if "nu.unif" not in pr:
            raise ValueError("error: nu.Unif must be specified")
        nu_Unif = _pair(pr["nu.unif"], "nu.Unif", increasing=True)

    st = _lower(starting)
    if "beta" in st:
        beta_starting = np.asarray(st["beta"], dtype=np.float64).reshape(-1)
        if beta_starting.size != p:
            raise ValueError("error: supplied beta.starting is the wrong "
                             "length")
    else:  # coefficients(lm(y ~ X - 1))
        beta_starting = np.linalg.lstsq(X_o, y_o, rcond=None)[0]
    for k in ("sigma.sq", "tau.sq", "phi") + (("nu",) if matern else ()):
        if k not in st:
            raise ValueError(f"error: {k} must be specified in starting "
                             "value list")
    tu = _lower(tuning)
    need_tu = ("phi",) if latent else ("sigma.sq", "tau.sq", "phi")
    for k in need_tu + (("nu",) if matern else ()):
        if k not in tu:
            raise ValueError(f"error: {k} must be specified in tuning value "
                             "list")

    def first(d, k):
        return float(np.asarray(d[k]).reshape(-1)[0])

    if neighbor_info is not None:
        nn_indx = np.asarray(neighbor_info["nn.indx"], dtype=np.int32)
        nn_indx_lu = np.asarray(neighbor_info["nn.indx.lu"], dtype=np.int32)
        nn_run_time = neighbor_info.get("nn.indx.run.time", 0.0)
    else:
        idx = _neighbors(coords_o, n_neighbors, search_type)
        nn_indx, nn_indx_lu = idx["nnIndx"], idx["nnIndxLU"]
        nn_run_time = idx["run.time"]

    n_rep, rep_indx, s_indx = 0, None, None
    if fit_rep:
        sub_sample, s_indx = _sub_sample(sub_sample, n_samples,
                                         int(np.floor(0.5 * n_samples)))
        rep_indx = np.zeros(n_samples, dtype=np.int32)
        rep_indx[s_indx - 1] = 1
        n_rep = int(len(s_indx))

    if latent:
        return _sNNGP_fit(
            y_o, X_o, coords_o, ord, p, n, n_neighbors, cov_model, matern,
            nn_indx, nn_indx_lu, nn_run_time, neighbor_info, sigma_sq_IG,
            tau_sq_IG, phi_Unif, nu_Unif, beta_starting, st, tu, first,
            n_samples, n_omp_threads, verbose, n_report, rng, seed,
            u_search_type, sweep, fit_rep, sub_sample, s_indx,
            return_neighbor_info, x_names, starting, priors, tuning)

    t0 = time.perf_counter()
    out = _nn.rNNGP(
        y_o, X_o, p, n, n_neighbors, coords_o, cov_model, nn_indx,
        nn_indx_lu, sigma_sq_IG, tau_sq_IG, phi_Unif, nu_Unif,
        beta_starting, first(st, "sigma.sq"), first(st, "tau.sq"),
        first(st, "phi"), first(st, "nu") if matern else 0.0,
        first(tu, "sigma.sq"), first(tu, "tau.sq"), first(tu, "phi"),
        first(tu, "nu") if matern else 0.0, n_samples, n_omp_threads,
        int(verbose), n_report, n_rep, rep_indx, rng=rng, seed=seed,
        precision=precision)
    run_time = time.perf_counter() - t0

    res = {"p.beta.samples": out["p.beta.samples"].T.copy(),
           "p.theta.samples": out["p.theta.samples"].T.copy(),
           "theta.names": ["sigma.sq", "tau.sq", "phi"]
                          + (["nu"] if matern else []),
           "x.names": list(x_names) if x_names is not None
                      else [f"x{j + 1}" for j in range(p)],
           "run.time": run_time, "accept": out.get("accept")}
    if return_neighbor_info:
        res["neighbor.info"] = {
            "n.neighbors": n_neighbors,
            "n.indx": n_indx_list(nn_indx, nn_indx_lu, n),
            "nn.indx": nn_indx, "nn.indx.lu": nn_indx_lu, "ord": ord,
            "nn.indx.run.time": nn_run_time}
    inv = np.argsort(ord, kind="stable")  # order(ord)
    res["coords"], res["y"], res["X"] = coords_o[inv], y_o[inv], X_o[inv]
    if fit_rep:
        bs = res["p.beta.samples"][s_indx - 1]
        res["y.hat.samples"] = res["X"] @ bs.T
        res["y.hat.quants"] = np.quantile(res["y.hat.samples"],
                                          [0.5, 0.05, 0.975], axis=1).T
        res["y.rep.samples"] = out["y.rep.samples"][inv]
        res["y.rep.quants"] = np.quantile(res["y.rep.samples"],
                                          [0.5, 0.05, 0.975], axis=1).T
        res["sub.sample"], res["s.indx"] = sub_sample, s_indx
    res.update({"n.neighbors": n_neighbors, "cov.model": cov_model,
                "cov.model.indx": COV_MODELS.index(cov_model),
                "starting": starting, "priors": priors, "tuning": tuning,
                "type": ["response", "gaussian"]})
    return res


def _normals(rng, k, gen):
    """k standard normals from the caller's source when it has one (R's
    rnorm(k) order), else from the numpy Generator gen."""
    if rng is not None and hasattr(rng, "norm"):
        return np.asarray(rng.norm(k), dtype=np.float64)
    return gen.standard_normal(k)


def _sNNGP_fit(y_o, X_o, coords_o, ord, p, n, m, cov_model, matern, nn_indx,
               nn_indx_lu, nn_run_time, neighbor_info, sigma_sq_IG, tau_sq_IG,
               phi_Unif, nu_Unif, beta_starting, st, tu, first, n_samples,
               n_omp_threads, verbose, n_report, rng, seed, u_search_type,
               sweep, fit_rep, sub_sample, s_indx, return_neighbor_info,
               x_names, starting, priors, tuning):
    """The latent branch of R/spNNGP.R: the neighbors-of-neighbors index
    (:332-347, or from neighbor.info :117-123), .Call("sNNGP") (:409-413),
    and the latent fitted values and replicates (:477-484)."""
    t0 = time.perf_counter()
    if neighbor_info is not None and "u.indx" in neighbor_info:
        u = {k: np.asarray(neighbor_info[k], dtype=np.int32)
             for k in ("u.indx", "u.indx.lu", "ui.indx")}
        u_run_time = neighbor_info.get("u.indx.run.time", 0.0)
    else:
        u = _nn.mkUIndx(n, m, nn_indx, nn_indx_lu, u_search_type)
        u_run_time = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = _nn.sNNGP(
        y_o, X_o, p, n, m, coords_o, cov_model, nn_indx, nn_indx_lu,
        u["u.indx"], u["u.indx.lu"], u["ui.indx"], sigma_sq_IG, tau_sq_IG,
        phi_Unif, nu_Unif, beta_starting, first(st, "sigma.sq"),
        first(st, "tau.sq"), first(st, "phi"), first(st, "nu") if matern
        else 0.0, first(tu, "sigma.sq") if "sigma.sq" in tu else 0.0,
        first(tu, "tau.sq") if "tau.sq" in tu else 0.0, first(tu, "phi"),
        first(tu, "nu") if matern else 0.0, n_samples, n_omp_threads,
        int(verbose), n_report, rng=rng, seed=seed, sweep=sweep)
    run_time = time.perf_counter() - t0
    res = {"p.beta.samples": out["p.beta.samples"].T.copy(),
           "p.theta.samples": out["p.theta.samples"].T.copy(),
           "theta.names": ["sigma.sq", "tau.sq", "phi"]
                          + (["nu"] if matern else []),
           "x.names": list(x_names) if x_names is not None
                      else [f"x{j + 1}" for j in range(p)],
           "run.time": run_time, "accept": out.get("accept")}
    if return_neighbor_info:
        res["neighbor.info"] = {
            "n.neighbors": m, "n.indx": n_indx_list(nn_indx, nn_indx_lu, n),
            "nn.indx": nn_indx, "nn.indx.lu": nn_indx_lu,
            "u.indx": u["u.indx"], "u.indx.lu": u["u.indx.lu"],
            "ui.indx": u["ui.indx"], "ord": ord,
            "nn.indx.run.time": nn_run_time, "u.indx.run.time": u_run_time}
    inv = np.argsort(ord, kind="stable")  # order(ord)
    res["coords"], res["y"], res["X"] = coords_o[inv], y_o[inv], X_o[inv]
    res["p.w.samples"] = out["p.w.samples"][inv]
    if fit_rep:
        bs = res["p.beta.samples"][s_indx - 1]
        yh = res["X"] @ bs.T + res["p.w.samples"][:, s_indx - 1]
        res["y.hat.samples"] = yh
        res["y.hat.quants"] = np.quantile(yh, [0.5, 0.05, 0.975], axis=1).T
        gen = np.random.default_rng(seed)
        tau = res["p.theta.samples"][s_indx - 1, 1]
        noise = np.column_stack([np.sqrt(t) * _normals(rng, n, gen)
                                 for t in tau])
        res["y.rep.samples"] = yh + noise
        res["y.rep.quants"] = np.quantile(res["y.rep.samples"],
                                          [0.5, 0.05, 0.975], axis=1).T
        res["sub.sample"], res["s.indx"] = sub_sample, s_indx
    res.update({"n.neighbors": m, "cov.model": cov_model,
                "cov.model.indx": COV_MODELS.index(cov_model),
                "starting": starting, "priors": priors, "tuning": tuning,
                "type": ["latent", "gaussian"], "sweep": sweep})
    return res


# ---------------------------------------------------------------------------
# spConjNNGP()
# ---------------------------------------------------------------------------
def spConjNNGP(y, X, coords, *, theta_alpha, sigma_sq_IG, n_neighbors=15,
               cov_model="exponential", k_fold=5, score_rule="crps",
               X_0=None, coords_0=None, n_omp_threads=1, search_type="cb",
               ord=None, return_neighbor_info=True, neighbor_info=None,
               fit_rep=False, n_samples=0, verbose=False, seed=1,
               ho_indx=None, x_names=None):
    """spConjNNGP(y ~ X - 1, coords, ...) of R/spConjNNGP.R.  theta_alpha:
    a dict {'phi', 'alpha'[, 'nu']} of equal-length vectors (a grid) or of
    scalars.  The k-fold search runs as one batched graph (CvGrid); the
    final fit and predictions use the cNNGP driver.  ho_indx (k x nho,
    0-based, in the ordered rows) overrides the random folds."""
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    coords = np.asarray(coords, dtype=np.float64)
    n, p = X.shape
    if coords.shape != (n, 2):
        raise ValueError("error: either coords has more than two columns or "
                         "the number of rows is different than data used "
                         "in the model formul")
    cov_model = _cov_name(cov_model)
    matern = cov_model == "matern"
    if fit_rep and n_samples <= 0:
        raise ValueError("error: if fit.rep=TRUE, specify the number of "
                         "fitted and replicated samples to generate using "
                         "the n.samples argument.")
    if neighbor_info is not None:
        ord = np.asarray(neighbor_info["ord"], dtype=np.int64)
        n_neighbors = int(neighbor_info["n.neighbors"])
    ord = _order(coords, ord)
    coords_o, X_o, y_o = coords[ord], X[ord], y[ord]

    ta = {str(k).lower(): np.atleast_1d(np.asarray(v, dtype=np.float64))
          for k, v in theta_alpha.items()}
    for k in ("phi", "alpha") + (("nu",) if matern else ()):
        if k not in ta:
            raise ValueError(f"error: column name {k} must be specified in "
                             "theta.alpha")
    cols = ["phi", "alpha"] + (["nu"] if matern else [])
    grid = np.column_stack([ta[k] for k in cols])  # g x nTheta
    g = grid.shape[0]
    single = g == 1
    if not single and (k_fold is None or k_fold < 2 or k_fold >= n):
        raise ValueError("error: k.fold must be greater than 1 and less "
                         "than n")
    ig = _pair(sigma_sq_IG, "sigma.sq.IG")

    if neighbor_info is not None:
        nn_indx = np.asarray(neighbor_info["nn.indx"], dtype=np.int32)
        nn_indx_lu = np.asarray(neighbor_info["nn.indx.lu"], dtype=np.int32)
        nn_run_time = neighbor_info.get("nn.indx.run.time", 0.0)
    else:
        idx = _neighbors(coords_o, n_neighbors, search_type)
        nn_indx, nn_indx_lu = idx["nnIndx"], idx["nnIndxLU"]
        nn_run_time = idx["run.time"]

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    res = {}
    best = grid[0]
    if not single:
        if score_rule not in ("rmspe", "crps"):
            raise ValueError(f"error: specified score.rule '{score_rule}' "
                             "is not a valid option; choose from rmspe, "
                             "crps.")
        nho = n // k_fold
        if ho_indx is None:  # sort(sample(1:n, floor(n/k.fold))) per fold
            ho_indx = np.array([np.sort(rng.choice(n, nho, replace=False))
                                for _ in range(k_fold)], dtype=np.int32)
        with _nn.CvGrid(y_o, X_o, coords_o, n_neighbors, ho_indx, ig,
                        cov_model, search_type=search_type) as cv:
            r = cv.run(grid)
        scores = r["scores"]  # g x 2: rmspe, crps
        best = grid[int(np.argmin(scores[:, 0 if score_rule == "rmspe"
                                          else 1]))]
        res["k.fold.scores"] = np.column_stack([grid, scores])
        res["k.fold.scores.names"] = cols + ["rmspe", "crps"]
        res["cv.graph.ms"] = r["graph_ms"]

    # final run with the chosen point (R/spConjNNGP.R:359-381)
    n0 = 0 if X_0 is None else np.asarray(X_0).shape[0]
    nn0 = None
    if n0:
        X_0 = np.asarray(X_0, dtype=np.float64)
        coords_0 = np.asarray(coords_0, dtype=np.float64)
        nn0, _ = _nn.nn2(coords_o, coords_0, n_neighbors)
    out = _nn.cNNGP(y_o, X_o, p, n, coords_o, best.reshape(-1, 1),
                    nn_indx, nn_indx_lu, X_0, coords_0, n0, nn0, 1,
                    n_neighbors, ig, cov_model, n_omp_threads, int(verbose))
    res["run.time"] = time.perf_counter() - t0
    ab = out["ab"][:, 0]
    res["theta.alpha"] = dict(zip(cols, best))
    res["ab"] = ab
    res["beta.hat"] = out["beta.hat"][:, 0]
    bBinv = out["bB.inv"][:, 0].reshape(p, p).T
    bBinv = np.tril(bBinv) + np.tril(bBinv, -1).T
    res["bB.inv"] = bBinv
    res["bb"] = out["bb"][:, 0]
    res["sigma.sq.hat"] = ab[1] / (ab[0] - 1)
    res["sigma.sq.var"] = ab[1] ** 2 / ((ab[0] - 1) ** 2 * (ab[0] - 2))
    res["beta.var"] = ab[1] * bBinv / (ab[0] - 1)
    if n0:
        res["y.0.hat"] = out["y.0.hat"][:, 0]
        res["y.0.hat.var"] = out["y.0.hat.var"][:, 0]
    res["x.names"] = (list(x_names) if x_names is not None
                      else [f"x{j + 1}" for j in range(p)])
    res.update({"sigma.sq.IG": ig, "n.neighbors": n_neighbors,
                "cov.model": cov_model,
                "cov.model.indx": COV_MODELS.index(cov_model),
                "search.type": search_type, "type": ["conjugate", "gaussian"]})
    if return_neighbor_info:
        res["neighbor.info"] = {
            "n.neighbors": n_neighbors,
            "n.indx": n_indx_list(nn_indx, nn_indx_lu, n),
            "nn.indx": nn_indx, "nn.indx.lu": nn_indx_lu, "ord": ord,
            "nn.indx.run.time": nn_run_time}
    if n_samples > 0:  # exact posterior samples (R/spConjNNGP.R:423-438)
        res["p.beta.theta.samples"] = _conj_samples(ab, bBinv, res["bb"],
                                                    best[1], n_samples, rng)
    res["coords"] = coords_o[np.argsort(ord, kind="stable")]
    res["X"] = X_o[np.argsort(ord, kind="stable")]
    res["y"] = y_o[np.argsort(ord, kind="stable")]
    if fit_rep:  # R/spConjNNGP.R:441-495
        samp = res["p.beta.theta.samples"]
        beta = samp[:, :p]
        res["y.hat.samples"] = res["X"] @ beta.T
        res["y.hat.quants"] = _quants(res["y.hat.samples"])
        res["y.rep.samples"] = _replicates(res, beta, _conj_theta(res, samp),
                                           verbose, 100, seed)
        res["y.rep.quants"] = _quants(res["y.rep.samples"])
        res["sub.sample"] = {"start": 1, "end": n_samples, "thin": 1}
        res["s.indx"] = np.arange(1, n_samples + 1)
    return res


def _conj_samples(ab, bBinv, bb, alpha, S, rng):
    """S exact posterior draws of (beta, sigma.sq, tau.sq) of the conjugate
    model (R/spConjNNGP.R:423-438, R/generics.R:54-61): sigma.sq ~ IG(a, b),
    beta ~ N(B^-1 b, sigma.sq B^-1), tau.sq = alpha sigma.sq."""
    p = len(bb)
    samp = np.zeros((S, p + 2))
    for i in range(S):
        s2 = 1.0 / rng.gamma(ab[0], 1.0 / ab[1])  # rigamma(1, a, b)
        L = np.linalg.cholesky(s2 * bBinv).T      # chol(V), upper
        beta = rng.normal(size=p) @ L + bBinv @ bb
        samp[i] = np.concatenate([beta, [s2, alpha * s2]])
    return samp


def _conj_theta(obj, samp):
    """theta (nTheta x S) of rNNGPReplicated for conjugate draws: sigma.sq,
    tau.sq, then phi (and nu) of the chosen grid point repeated
    (R/spConjNNGP.R:442-447)."""
    S = samp.shape[0]
    ta = obj["theta.alpha"]
    rows = [samp[:, -2], samp[:, -1], np.full(S, ta["phi"])]
    if obj["cov.model"] == "matern":
        rows.append(np.full(S, ta["nu"]))
    return np.vstack(rows)


def _quants(a):
    return np.quantile(a, [0.5, 0.05, 0.975], axis=1).T


def _replicates(obj, beta, theta, verbose, n_report, seed):
    """y.rep of rNNGPReplicated for the object's data (R/generics.R:223-261),
    in the caller's row order.  beta S x p, theta nTheta x S."""
    ni = obj.get("neighbor.info")
    if ni is None:
        raise ValueError("error: the neighbor.info is not in the model object, "
                         "call the fitting function again with "
                         "return_neighbor_info=True.")
    ord = np.asarray(ni["ord"])
    X_o = np.asarray(obj["X"], np.float64)[ord]
    c_o = np.asarray(obj["coords"], np.float64)[ord]
    n, p = X_o.shape
    beta = np.asarray(beta, np.float64)
    out = _nn.rNNGPReplicated(X_o, p, n, obj["n.neighbors"], c_o,
                              obj["cov.model"], ni["nn.indx"],
                              ni["nn.indx.lu"], beta, theta, beta.shape[0],
                              verbose=int(verbose), nReport=n_report,
                              seed=seed)
    inv = np.argsort(ord, kind="stable")
    return np.asarray(out["y.rep.samples"])[inv]


# ---------------------------------------------------------------------------
# fitted() (R/generics.R:1-268)
# ---------------------------------------------------------------------------
def fitted(obj, sub_sample=None, *, n_omp_threads=1, verbose=False,
           n_report=100, seed=1):
    """fitted.NNGP: fitted values and replicates with their median and 95%
    quantiles.  Latent: X beta + w and y.rep = that + tau.sq noise.
    Response and conjugate: fitted values X beta and replicates y.rep
    (rNNGPReplicated on the device) with their
    median and 95% quantiles.  Conjugate: sub_sample is the number of exact
    posterior draws (stored ones are returned when it matches or is None).
    Response: sub_sample = {'start', 'end', 'thin'} over the stored chain
    (1-based; default start floor(n.samples / 2)); without it the object's
    fit.rep results are returned.  Parameter draws use numpy's Generator
    and the replicate normals the library's counter-based stream, both
    seeded by `seed` (R's RNG is not available)."""
    del n_omp_threads
    keys = ("y.hat.quants", "y.hat.samples", "y.rep.quants", "y.rep.samples")
    if obj["type"][0] == "conjugate":
        if sub_sample is not None and not (np.ndim(sub_sample) == 0
                                           and sub_sample > 0):
            raise ValueError("error: for a conjugate model, the sub.sample "
                             "argument should be in integer value indicating "
                             "the number of fitted and replicated samples to "
                             "collect.")
        if "y.rep.samples" in obj and (
                sub_sample is None
                or int(sub_sample) == obj["y.rep.samples"].shape[1]):
            return {k: obj[k] for k in keys}
        if sub_sample is None:
            raise ValueError("error: for an conjugate model created with "
                             "fit.rep=FALSE, the sub.sample argument should be "
                             "provided to indicate the number of fitted and "
                             "replicated samples to collect.")
        S = int(sub_sample)
        p = len(obj["bb"])
        samp = _conj_samples(obj["ab"], obj["bB.inv"], obj["bb"],
                             obj["theta.alpha"]["alpha"], S,
                             np.random.default_rng(seed))
        beta = samp[:, :p]
        out = {"p.beta.theta.samples": samp,
               "y.hat.samples": np.asarray(obj["X"]) @ beta.T}
        out["y.hat.quants"] = _quants(out["y.hat.samples"])
        out["y.rep.samples"] = _replicates(obj, beta, _conj_theta(obj, samp),
                                           verbose, n_report, seed)
        out["y.rep.quants"] = _quants(out["y.rep.samples"])
        out["sub.sample"] = {"start": 1, "end": S, "thin": 1}
        out["s.indx"] = np.arange(1, S + 1)
        return out
    if sub_sample is None and "s.indx" in obj:
        return {k: obj[k] for k in keys + ("sub.sample", "s.indx")}
    n_samples = np.asarray(obj["p.beta.samples"]).shape[0]
    ss, s_indx = _sub_sample(sub_sample, n_samples,
                             int(np.floor(0.5 * n_samples)))
    if obj["type"][0] == "latent":
        # X beta + w per kept sample, replicates add tau.sq noise
        # (R/generics.R:187-200)
        beta = np.asarray(obj["p.beta.samples"])[s_indx - 1]
        X = np.asarray(obj["X"])
        yh = X @ beta.T + np.asarray(obj["p.w.samples"])[:, s_indx - 1]
        gen = np.random.default_rng(seed)
        tau = np.asarray(obj["p.theta.samples"])[s_indx - 1, 1]
        out = {"y.hat.samples": yh, "y.hat.quants": _quants(yh),
               "y.rep.samples": yh + np.column_stack(
                   [np.sqrt(t) * gen.standard_normal(X.shape[0])
                    for t in tau])}
        out["y.rep.quants"] = _quants(out["y.rep.samples"])
        out["sub.sample"], out["s.indx"] = ss, s_indx
        return out
    beta = np.asarray(obj["p.beta.samples"])[s_indx - 1]
    theta = np.asarray(obj["p.theta.samples"])[s_indx - 1].T.copy()
    out = {"y.hat.samples": np.asarray(obj["X"]) @ beta.T}
    out["y.hat.quants"] = _quants(out["y.hat.samples"])
    out["y.rep.samples"] = _replicates(obj, beta, theta, verbose, n_report,
                                       seed)
    out["y.rep.quants"] = _quants(out["y.rep.samples"])
    out["sub.sample"], out["s.indx"] = ss, s_indx
    return out


# ---------------------------------------------------------------------------
# predict() (R/generics.R:583-675)
# ---------------------------------------------------------------------------
def predict(obj, X_0, coords_0, *, sub_sample=None, n_omp_threads=1,
            verbose=False, n_report=100, rng=None, seed=1, server=None,
            request_seed=None):
    """predict.NNGP for a response fit (posterior predictive samples p.y.0,
    q x n.samples), a latent fit (p.y.0 and the latent p.w.0) or a
    conjugate fit (y.0.hat, y.0.hat.var).

    For a response fit, `server` (a spnngp_amd.Server built on the same
    fitted model) answers the request from HBM instead, with the counter-
    based draws keyed by request_seed."""
    X_0 = np.asarray(X_0, dtype=np.float64)
    coords_0 = np.asarray(coords_0, dtype=np.float64)
    X, y, coords = obj["X"], obj["y"], obj["coords"]
    n, p = X.shape
    if X_0.ndim != 2 or X_0.shape[1] != p:
        raise ValueError(f"error: X.0 must have  {p}  columns\n")
    if coords_0.ndim != 2 or coords_0.shape[1] != 2:
        raise ValueError("error: coords.0 must have two columns\n")
    q = X_0.shape[0]
    m = obj["n.neighbors"]
    if obj["type"][0] == "conjugate":
        ta = obj["theta.alpha"]
        r = spConjNNGP(y, X, coords, theta_alpha={k: [v] for k, v in
                                                  ta.items()},
                       sigma_sq_IG=obj["sigma.sq.IG"], n_neighbors=m,
                       cov_model=obj["cov.model"], X_0=X_0,
                       coords_0=coords_0, search_type=obj.get("search.type",
                                                              "cb"),
                       return_neighbor_info=False)
        return {"y.0.hat": r["y.0.hat"], "y.0.hat.var": r["y.0.hat.var"]}
    S_all = obj["p.beta.samples"].shape[0]
    sub_sample, s_indx = _sub_sample(sub_sample, S_all, 1)
    beta = obj["p.beta.samples"][s_indx - 1].T
    theta = obj["p.theta.samples"][s_indx - 1].T
    S = beta.shape[1]
    t0 = time.perf_counter()
    if obj["type"][0] == "latent":
        # src/sNNGPPredict.cpp via R/generics.R:600-672: the kept w samples
        # at the fitted sites condition the new ones
        if server is not None:
            raise ValueError("the kriging server serves response fits; a "
                             "latent fit is predicted without one")
        nn0, _ = _nn.nn2(coords, coords_0, m)
        w = np.asarray(obj["p.w.samples"])[:, s_indx - 1]
        out = _nn.sNNGPPredict(X, y, coords, n, p, m, X_0, coords_0, q, nn0,
                               beta, theta, w, S, obj["cov.model"],
                               nThreads=n_omp_threads, verbose=int(verbose),
                               nReport=n_report, rng=rng, seed=seed)
        return {"p.y.0": out["p.y.0"], "p.w.0": out["p.w.0"],
                "run.time": time.perf_counter() - t0,
                "sub.sample": sub_sample}
    if server is not None:
        out = server.predict(X_0, coords_0,
                             seed if request_seed is None else request_seed)
    else:
        nn0, _ = _nn.nn2(coords, coords_0, m)
        out = _nn.rNNGPPredict(X, y, coords, n, p, m, X_0, coords_0, q, nn0,
                               beta, theta, S, obj["cov.model"],
                               n_omp_threads, int(verbose), n_report,
                               rng=rng, seed=seed)
    return {"p.y.0": out["p.y.0"], "run.time": time.perf_counter() - t0,
            "sub.sample": sub_sample}


# ---------------------------------------------------------------------------
# fitted-model format: save / load a fit, and serve it
# ---------------------------------------------------------------------------
# The arrays of a fit object (the R list's elements with R's names:
# p.beta.samples S x p, p.theta.samples S x nTheta, X, y, coords in the
# caller's row order, neighbor.info's nn.indx / nn.indx.lu / ord, ...) go
# into one .npz as plain numeric arrays; every non-array element (names,
# scalars, nested dicts of scalars) goes into a JSON document stored as a
# string array.  Loading needs no pickle (numpy.load with
# allow_pickle=False), so a serving process can read fits it did not make.
_FIT_FORMAT = "spnngp_amd.fit/1"


def _flatten_fit(obj):
    arrays, meta = {}, {}
    for k, v in obj.items():
        if k == "neighbor.info" and isinstance(v, dict):
            ni = {}
            for kk, vv in v.items():
                if kk == "n.indx":
                    continue  # rebuilt from nn.indx / nn.indx.lu
                if isinstance(vv, np.ndarray):
                    arrays["neighbor.info/" + kk] = vv
                else:
                    ni[kk] = vv
            meta["neighbor.info"] = ni
        elif isinstance(v, np.ndarray):
            if v.dtype == object:
                raise ValueError(f"fit element {k!r} is not a numeric array")
            arrays[k] = v
        else:
            meta[k] = v
    return arrays, meta


def _json_default(v):
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, np.ndarray):
        return v.tolist()
    raise TypeError(f"cannot store {type(v).__name__} in a fit file")


def save_fit(obj, path):
    """Write a spNNGP / spConjNNGP fit (the dict these functions return) to
    `path` (.npz): numeric arrays as they are, everything else as JSON."""
    import json
    arrays, meta = _flatten_fit(obj)
    doc = json.dumps({"format": _FIT_FORMAT, "meta": meta},
                     default=_json_default)
    np.savez(path, __meta__=np.array(doc), **arrays)


def load_fit(path):
    """Read a fit written by save_fit (no pickle is loaded)."""
    import json
    with np.load(path, allow_pickle=False) as f:
        doc = json.loads(str(f["__meta__"]))
        if doc.get("format") != _FIT_FORMAT:
            raise ValueError(f"{path}: not a {_FIT_FORMAT} file")
        obj = dict(doc["meta"])
        ni = obj.pop("neighbor.info", None)
        for k in f.files:
            if k == "__meta__":
                continue
            if k.startswith("neighbor.info/"):
                ni = {} if ni is None else ni
                ni[k.split("/", 1)[1]] = f[k]
            else:
                obj[k] = f[k]
    if ni is not None:
        if "nn.indx" in ni and "nn.indx.lu" in ni:
            n = len(ni["nn.indx.lu"]) // 2
            ni["n.indx"] = n_indx_list(ni["nn.indx"], ni["nn.indx.lu"], n)
        obj["neighbor.info"] = ni
    return obj


def serve(obj, sub_sample=None, **server_opts):
    """A resident kriging server (spnngp_amd.Server) for a response fit:
    the fitted rows, their neighbor index and the posterior samples
    sub_sample = {'start', 'end', 'thin'} (1-based, predict()'s default:
    all samples) stay in HBM.  predict(obj, ..., server=...) then answers
    from it; server_opts go to Server (device_bytes, host_bytes,
    batch_sites, depth, device)."""
    if obj["type"][0] != "response":
        raise ValueError("serve() takes a response fit (spNNGP)")
    S_all = np.asarray(obj["p.beta.samples"]).shape[0]
    _, s_indx = _sub_sample(sub_sample, S_all, 1)
    beta = np.asarray(obj["p.beta.samples"])[s_indx - 1].T
    theta = np.asarray(obj["p.theta.samples"])[s_indx - 1].T
    return _nn.Server(obj["X"], obj["y"], obj["coords"], obj["n.neighbors"],
                      beta, theta, obj["cov.model"], **server_opts)
