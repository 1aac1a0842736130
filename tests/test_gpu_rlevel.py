"""The R-level front ends end to end (spNNGP, spConjNNGP, predict), each
against the oracle run on the same ordered data with the same draws."""
import numpy as np
import pytest

RTOL = 1e-10


def max_rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def gp_data(n, seed):
    """The CPU-runnable response case: U[0,1]^2, X = [1, N(0,1)],
    beta = (1, 5), w an exact exponential GP (sigma.sq 5, phi 6), tau.sq 1."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(size=(n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    D = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    L = np.linalg.cholesky(5 * np.exp(-6 * D) + 1e-8 * np.eye(n))
    y = X @ [1.0, 5.0] + L @ rng.normal(size=n) + rng.normal(size=n)
    return c, X, y


@pytest.mark.gpu
def test_spNNGP_matches_oracle_chain(oracle, nngp):
    print("spNNGP front end", flush=True)
    n, m, S = 600, 10, 150
    c, X, y = gp_data(n, 1)
    rng = np.random.default_rng(2)
    zs, us = rng.normal(size=S * 8), rng.uniform(size=S)
    priors = {"sigma.sq.IG": [2, 5], "tau.sq.IG": [2, 1],
              "phi.Unif": [3, 300]}
    starting = {"sigma.sq": 5, "tau.sq": 1, "phi": 6}
    tuning = {"sigma.sq": 0.5, "tau.sq": 0.5, "phi": 0.5}
    fit = nngp.spNNGP(y, X, c, starting=starting, tuning=tuning,
                      priors=priors, n_samples=S, n_neighbors=m,
                      rng=oracle.TapeRng(zs, us), return_neighbor_info=True)
    # the oracle on the x-ordered data with the OLS start
    o = np.argsort(c[:, 0], kind="stable")
    co, Xo, yo = c[o], X[o], y[o]
    nn = oracle.mkNNIndxCB(co, m)
    assert np.array_equal(fit["neighbor.info"]["nn.indx"], nn["nnIndx"])
    b0 = np.linalg.lstsq(Xo, yo, rcond=None)[0]
    ref = oracle.rNNGP(yo, Xo, 2, n, m, co.T.reshape(-1).copy(), 0,
                       nn["nnIndx"], nn["nnIndxLU"], sigmaSqIG=[2, 5],
                       tauSqIG=[2, 1], phiUnif=[3, 300], nuUnif=None,
                       betaStarting=b0, sigmaSqStarting=5.0,
                       tauSqStarting=1.0, phiStarting=6.0, nuStarting=0.0,
                       sigmaSqTuning=0.5, tauSqTuning=0.5, phiTuning=0.5,
                       nuTuning=0.0, nSamples=S,
                       rng=oracle.TapeRng(zs, us))
    assert max_rel(fit["p.beta.samples"], ref["p.beta.samples"].T) < RTOL
    assert max_rel(fit["p.theta.samples"], ref["p.theta.samples"].T) < RTOL
    # results come back in the caller's order
    assert np.array_equal(fit["coords"], c) and np.array_equal(fit["y"], y)


@pytest.mark.gpu
def test_predict_response_matches_oracle(oracle, nngp):
    print("predict front end", flush=True)
    n, m, S, q = 500, 8, 60, 40
    c, X, y = gp_data(n, 3)
    rng = np.random.default_rng(4)
    fit = nngp.spNNGP(y, X, c, starting={"sigma.sq": 5, "tau.sq": 1,
                                         "phi": 6},
                      tuning={"sigma.sq": 0.3, "tau.sq": 0.3, "phi": 0.3},
                      priors={"sigma.sq.IG": [2, 5], "tau.sq.IG": [2, 1],
                              "phi.Unif": [3, 300]},
                      n_samples=S, n_neighbors=m, seed=9)
    c0 = rng.uniform(size=(q, 2))
    X0 = np.column_stack([np.ones(q), rng.normal(size=q)])
    zs = rng.normal(size=q * S)
    got = nngp.predict(fit, X0, c0, sub_sample={"start": 11, "thin": 2},
                       rng=oracle.TapeRng(zs, np.zeros(1)))["p.y.0"]
    s_indx = np.arange(11, S + 1, 2) - 1
    beta = fit["p.beta.samples"][s_indx].T
    th = fit["p.theta.samples"][s_indx].T
    Sq = len(s_indx)
    nn0, _ = oracle.nn2(c, c0, m)
    ref = oracle.rNNGPPredict(X.T.reshape(-1).copy(), y,
                              c.T.reshape(-1).copy(), n, 2, m,
                              X0.T.reshape(-1).copy(),
                              c0.T.reshape(-1).copy(), q,
                              nn0.T.reshape(-1).copy(),
                              beta.T.reshape(-1).copy(),
                              th.T.reshape(-1).copy(), Sq, 0, zs[:q * Sq])
    assert got.shape == (q, Sq)
    assert max_rel(got, ref) < RTOL


@pytest.mark.gpu
def test_spConjNNGP_kfold_and_final_fit(oracle, nngp):
    print("spConjNNGP front end", flush=True)
    n, m, k = 900, 10, 3
    c, X, y = gp_data(n, 5)
    rng = np.random.default_rng(6)
    nho = n // k
    ho = np.array([np.sort(rng.choice(n, nho, replace=False))
                   for _ in range(k)], dtype=np.int32)
    grid = {"phi": [3.0, 3.0, 9.0, 9.0], "alpha": [0.1, 0.5, 0.1, 0.5]}
    c0 = rng.uniform(size=(25, 2))
    X0 = np.column_stack([np.ones(25), rng.normal(size=25)])
    fit = nngp.spConjNNGP(y, X, c, theta_alpha=grid, sigma_sq_IG=[2, 5],
                          n_neighbors=m, k_fold=k, X_0=X0, coords_0=c0,
                          ho_indx=ho, n_samples=50)
    o = np.argsort(c[:, 0], kind="stable")
    co, Xo, yo = c[o], X[o], y[o]
    ta = np.column_stack([grid["phi"], grid["alpha"]])
    want = np.zeros((4, 2))
    for f in range(k):
        keep = np.ones(n, bool)
        keep[ho[f]] = False
        nn = oracle.mkNNIndxCB(co[keep], m)
        nn0, _ = oracle.nn2(co[keep], co[ho[f]], m)
        r = oracle.cNNGP(yo[keep], Xo[keep].T.reshape(-1).copy(), 2,
                         int(keep.sum()), co[keep].T.reshape(-1).copy(),
                         ta.reshape(-1).copy(), nn["nnIndx"], nn["nnIndxLU"],
                         Xo[ho[f]].T.reshape(-1).copy(),
                         co[ho[f]].T.reshape(-1).copy(), nho,
                         nn0.T.reshape(-1).copy(), 4, m, [2, 5], 0)
        for j in range(4):
            want[j, 0] += nngp.rmspe(yo[ho[f]], r["y.0.hat"][:, j]) / k
            want[j, 1] += nngp.crps(yo[ho[f]], r["y.0.hat"][:, j],
                                    r["y.0.hat.var"][:, j]) / k
    assert max_rel(fit["k.fold.scores"][:, 2:], want) < RTOL
    best = ta[np.argmin(want[:, 1])]
    assert fit["theta.alpha"]["phi"] == best[0]
    assert fit["theta.alpha"]["alpha"] == best[1]
    nn = oracle.mkNNIndxCB(co, m)
    nn0, _ = oracle.nn2(co, c0, m)
    r = oracle.cNNGP(yo, Xo.T.reshape(-1).copy(), 2, n,
                     co.T.reshape(-1).copy(), best.copy(), nn["nnIndx"],
                     nn["nnIndxLU"], X0.T.reshape(-1).copy(),
                     c0.T.reshape(-1).copy(), 25, nn0.T.reshape(-1).copy(),
                     1, m, [2, 5], 0)
    assert max_rel(fit["beta.hat"], r["beta.hat"][:, 0]) < RTOL
    assert max_rel(fit["y.0.hat"], r["y.0.hat"][:, 0]) < RTOL
    assert max_rel(fit["y.0.hat.var"], r["y.0.hat.var"][:, 0]) < RTOL
    a, b = r["ab"][:, 0]
    assert fit["sigma.sq.hat"] == pytest.approx(b / (a - 1), rel=RTOL)
    assert fit["p.beta.theta.samples"].shape == (50, 4)


def oracle_replicates(oracle, fit, beta, theta, seed):
    """The oracle's rNNGPReplicated on the object's x-ordered data with the
    library's counter-based normals for `seed`, back in the caller's order."""
    ni = fit["neighbor.info"]
    o = np.asarray(ni["ord"])
    Xo, co = np.asarray(fit["X"])[o], np.asarray(fit["coords"])[o]
    n, p = Xo.shape
    S = beta.shape[0]
    z = np.array([oracle.lib().or_philox_normal(seed, i, s)
                  for s in range(S) for i in range(n)])
    ref = oracle.rNNGPReplicated(Xo.T.reshape(-1).copy(), p, n,
                                 fit["n.neighbors"], co.T.reshape(-1).copy(),
                                 fit["cov.model.indx"], ni["nn.indx"],
                                 ni["nn.indx.lu"], beta.T.reshape(-1).copy(),
                                 theta.T.reshape(-1).copy(), S,
                                 oracle.TapeRng(z, []))
    return ref[np.argsort(o, kind="stable")]


@pytest.mark.gpu
def test_spConjNNGP_fit_rep_and_fitted(oracle, nngp):
    """fit.rep of the conjugate model (R/spConjNNGP.R:441-495) and
    fitted() (R/generics.R:17-150): X beta of the exact draws, and the
    replicates of rNNGPReplicated on the chosen (phi, alpha)."""
    print("spConjNNGP fit.rep / fitted", flush=True)
    n, m, S, seed = 500, 8, 30, 4
    c, X, y = gp_data(n, 3)
    fit = nngp.spConjNNGP(y, X, c, theta_alpha={"phi": 6.0, "alpha": 0.2},
                          sigma_sq_IG=[2, 5], n_neighbors=m, fit_rep=True,
                          n_samples=S, seed=seed)
    samp = fit["p.beta.theta.samples"]
    assert fit["y.rep.samples"].shape == (n, S)
    assert max_rel(fit["y.hat.samples"], X @ samp[:, :2].T) < RTOL
    theta = np.vstack([samp[:, 2], samp[:, 3], np.full(S, 6.0)])
    ref = oracle_replicates(oracle, fit, samp[:, :2], theta, seed)
    assert max_rel(fit["y.rep.samples"], ref) < RTOL
    q = fit["y.rep.quants"]
    assert q.shape == (n, 3) and np.all(q[:, 1] <= q[:, 0])
    assert np.all(q[:, 0] <= q[:, 2])
    # fitted(): the stored replicates, or new draws on request
    same = nngp.fitted(fit)
    assert same["y.rep.samples"] is fit["y.rep.samples"]
    new = nngp.fitted(fit, 12, seed=9)
    assert new["y.rep.samples"].shape == (n, 12)
    s2 = new["p.beta.theta.samples"][:, 2]
    th = np.vstack([s2, 0.2 * s2, np.full(12, 6.0)])
    ref = oracle_replicates(oracle, fit, new["p.beta.theta.samples"][:, :2],
                            th, 9)
    assert max_rel(new["y.rep.samples"], ref) < RTOL


@pytest.mark.gpu
def test_fitted_response_model(oracle, nngp):
    """fitted() of a response fit (R/generics.R:152-268): X beta and the
    replicates of the chain's samples start..end by thin."""
    print("fitted response", flush=True)
    n, m, S, seed = 400, 6, 40, 2
    c, X, y = gp_data(n, 5)
    fit = nngp.spNNGP(y, X, c, starting={"sigma.sq": 5, "tau.sq": 1, "phi": 6},
                      tuning={"sigma.sq": 0.3, "tau.sq": 0.3, "phi": 0.3},
                      priors={"sigma.sq.IG": [2, 5], "tau.sq.IG": [2, 1],
                              "phi.Unif": [3, 300]},
                      n_samples=S, n_neighbors=m, return_neighbor_info=True)
    out = nngp.fitted(fit, {"start": 10, "end": 40, "thin": 5}, seed=seed)
    s_indx = np.arange(10, 41, 5)
    assert np.array_equal(out["s.indx"], s_indx)
    beta = fit["p.beta.samples"][s_indx - 1]
    theta = fit["p.theta.samples"][s_indx - 1].T
    assert max_rel(out["y.hat.samples"], X @ beta.T) < RTOL
    ref = oracle_replicates(oracle, fit, beta, theta, seed)
    assert max_rel(out["y.rep.samples"], ref) < RTOL


@pytest.mark.gpu
def test_saved_fit_serves_the_same_predictions(nngp, tmp_path):
    """A fit written by save_fit and read back by another reader serves
    bit-identical predictive draws (same request seed) to the original."""
    print("fit file -> server", flush=True)
    n, m, S = 500, 8, 60
    c, X, y = gp_data(n, 7)
    fit = nngp.spNNGP(y, X, c, starting={"sigma.sq": 5, "tau.sq": 1, "phi": 6},
                      tuning={"sigma.sq": 0.3, "tau.sq": 0.3, "phi": 0.3},
                      priors={"sigma.sq.IG": [2, 5], "tau.sq.IG": [2, 1],
                              "phi.Unif": [3, 300]},
                      n_samples=S, n_neighbors=m, return_neighbor_info=True)
    nngp.save_fit(fit, tmp_path / "fit.npz")
    back = nngp.load_fit(tmp_path / "fit.npz")
    rng = np.random.default_rng(8)
    c0 = rng.uniform(size=(150, 2))
    X0 = np.column_stack([np.ones(150), rng.normal(size=150)])
    ss = {"start": 21, "end": 60, "thin": 2}
    with nngp.serve(fit, ss) as s1, nngp.serve(back, ss) as s2:
        a = nngp.predict(fit, X0, c0, sub_sample=ss, server=s1,
                         request_seed=5)["p.y.0"]
        b = nngp.predict(back, X0, c0, sub_sample=ss, server=s2,
                         request_seed=5)["p.y.0"]
    assert a.shape == (150, 20)
    assert np.array_equal(a, b)


@pytest.mark.gpu
def test_spNNGP_chain_from_r_set_seed_stream(oracle, nngp):
    """With the library's R generator (set.seed(2024) stream) the chain is
    the oracle's chain driven by the same stream."""
    print("spNNGP with R's stream", flush=True)
    n, m, S = 500, 8, 120
    c, X, y = gp_data(n, 11)
    kw = dict(starting={"sigma.sq": 5, "tau.sq": 1, "phi": 6},
              tuning={"sigma.sq": 0.4, "tau.sq": 0.4, "phi": 0.4},
              priors={"sigma.sq.IG": [2, 5], "tau.sq.IG": [2, 1],
                      "phi.Unif": [3, 300]})
    fit = nngp.spNNGP(y, X, c, n_samples=S, n_neighbors=m,
                      rng=nngp.RRng(2024), **kw)
    o = np.argsort(c[:, 0], kind="stable")
    co, Xo, yo = c[o], X[o], y[o]
    nn_ = oracle.mkNNIndxCB(co, m)
    b0 = np.linalg.lstsq(Xo, yo, rcond=None)[0]
    ref = oracle.rNNGP(yo, Xo, 2, n, m, co.T.reshape(-1).copy(), 0,
                       nn_["nnIndx"], nn_["nnIndxLU"], sigmaSqIG=[2, 5],
                       tauSqIG=[2, 1], phiUnif=[3, 300], nuUnif=None,
                       betaStarting=b0, sigmaSqStarting=5.0,
                       tauSqStarting=1.0, phiStarting=6.0, nuStarting=0.0,
                       sigmaSqTuning=0.4, tauSqTuning=0.4, phiTuning=0.4,
                       nuTuning=0.0, nSamples=S, rng=nngp.RRng(2024))
    assert max_rel(fit["p.beta.samples"], ref["p.beta.samples"].T) < RTOL
    assert max_rel(fit["p.theta.samples"], ref["p.theta.samples"].T) < RTOL


def r_get_n_indx(i, m):
    """get.n.indx of R/util.R:257-270, 1-based as in R: the positions of
    site i's neighbors in nn.indx (None for the first site)."""
    i = i - 1
    if i == 0:
        return None
    if i < m:
        start = i / 2 * (i - 1)
        return np.arange(start + 1, start + 1 + i).astype(int)
    start = m / 2 * (m - 1) + (i - m) * m
    return np.arange(start + 1, start + 1 + m).astype(int)


@pytest.mark.gpu
def test_neighbor_info_and_sample_layout_match_the_reference(oracle, nngp,
                                                             tmp_path):
    """The fit's list layout of R/spNNGP.R:439-456: p.theta.samples and
    p.beta.samples transposed to n.samples rows (mcmc(t(...))) with the
    reference's column names; neighbor.info with n.neighbors, the
    mk.n.indx.list of R/util.R:272-279 (NA, then 1-based neighbor sets
    through get.n.indx), nn.indx / nn.indx.lu of mkNNIndxCB on the x-ordered
    sites, and ord = order(coords[, 1]); coords, y and X back in the
    caller's order.  save_fit / load_fit keep every element."""
    print("neighbor.info layout", flush=True)
    n, m, S = 500, 8, 30
    c, X, y = gp_data(n, 31)
    c[::50, 0] = c[1::50, 0]  # tied x values: order() is stable
    rng = np.random.default_rng(31)
    zs, us = rng.normal(size=S * 8), rng.uniform(size=S)
    fit = nngp.spNNGP(y, X, c, starting={"sigma.sq": 5, "tau.sq": 1,
                                         "phi": 6},
                      tuning={"sigma.sq": 0.5, "tau.sq": 0.5, "phi": 0.5},
                      priors={"sigma.sq.IG": [2, 5], "tau.sq.IG": [2, 1],
                              "phi.Unif": [3, 300]},
                      n_samples=S, n_neighbors=m,
                      rng=oracle.TapeRng(zs, us), return_neighbor_info=True)
    ord_r = np.argsort(c[:, 0], kind="stable")  # order(coords[,1])
    co, Xo, yo = c[ord_r], X[ord_r], y[ord_r]
    nn = oracle.mkNNIndxCB(co, m)
    ni = fit["neighbor.info"]
    assert ni["n.neighbors"] == m
    assert np.array_equal(ni["ord"], ord_r)
    assert np.array_equal(ni["nn.indx"], nn["nnIndx"])
    assert np.array_equal(ni["nn.indx.lu"], nn["nnIndxLU"])
    assert len(ni["n.indx"]) == n and ni["n.indx"][0] is None
    for i in range(2, n + 1):
        pos = r_get_n_indx(i, m)
        assert np.array_equal(ni["n.indx"][i - 1], nn["nnIndx"][pos - 1] + 1)
    ref = oracle.rNNGP(yo, Xo, 2, n, m, co.T.reshape(-1).copy(), 0,
                       nn["nnIndx"], nn["nnIndxLU"], sigmaSqIG=[2, 5],
                       tauSqIG=[2, 1], phiUnif=[3, 300], nuUnif=None,
                       betaStarting=np.linalg.lstsq(Xo, yo, rcond=None)[0],
                       sigmaSqStarting=5.0, tauSqStarting=1.0,
                       phiStarting=6.0, nuStarting=0.0, sigmaSqTuning=0.5,
                       tauSqTuning=0.5, phiTuning=0.5, nuTuning=0.0,
                       nSamples=S, rng=oracle.TapeRng(zs, us))
    assert fit["p.theta.samples"].shape == (S, 3)
    assert fit["p.beta.samples"].shape == (S, 2)
    assert fit["theta.names"] == ["sigma.sq", "tau.sq", "phi"]
    assert max_rel(fit["p.theta.samples"], ref["p.theta.samples"].T) < RTOL
    assert max_rel(fit["p.beta.samples"], ref["p.beta.samples"].T) < RTOL
    assert np.array_equal(fit["coords"], c) and np.array_equal(fit["y"], y)
    assert np.array_equal(fit["X"], X)
    path = tmp_path / "fit.npz"
    nngp.save_fit(fit, path)
    back = nngp.load_fit(path)
    for k in ("p.theta.samples", "p.beta.samples", "coords", "y", "X"):
        assert np.array_equal(back[k], fit[k]), k
    bni = back["neighbor.info"]
    for k in ("nn.indx", "nn.indx.lu", "ord"):
        assert np.array_equal(bni[k], ni[k]), k
    assert bni["n.indx"][0] is None
    assert all(np.array_equal(a, b) for a, b in
               zip(bni["n.indx"][1:], ni["n.indx"][1:]))
