"""Conjugate k-fold grid search (nngp_cvgrid_*): every fold x grid point
against the oracle's cNNGP on that fold (beta.hat, a, b, (X'V^-1X)^-1,
held-out predictions and variances), and the scores against the CRPS /
RMSPE of R/util.R:247-255 computed from the oracle's predictions."""
import numpy as np
import pytest
import scipy.stats as ss

RTOL = 1e-10


def max_rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def r_crps(y, yhat, var):
    sd = np.sqrt(var)
    z = (y - yhat) / sd
    return np.mean(-sd * (1 / np.sqrt(np.pi) - 2 * ss.norm.pdf(z)
                          - z * (2 * ss.norm.cdf(z) - 1)))


def r_rmspe(y, yhat):
    return np.sqrt(np.mean((y - yhat) ** 2))


def data(n, p, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(size=(n, 2))
    c = c[np.argsort(c[:, 0], kind="stable")]
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))])
    w = np.sin(6 * c[:, 0]) * np.cos(5 * c[:, 1])
    y = X @ np.linspace(1, -1, p) + w + 0.3 * rng.normal(size=n)
    return c, X, y


def folds(n, k, nho, seed):
    rng = np.random.default_rng(seed)
    return np.array([np.sort(rng.choice(n, nho, replace=False))
                     for _ in range(k)], dtype=np.int32)


@pytest.mark.gpu
@pytest.mark.parametrize("cov,nu,m", [(0, None, 10), (2, 1.5, 12),
                                      (2, 0.9, 8), (1, None, 15),
                                      (3, None, 6), (0, None, 20)])
def test_cvgrid_matches_oracle_per_fold(oracle, nngp, cov, nu, m):
    print(f"cvgrid cov={cov} nu={nu} m={m}", flush=True)
    n, p, k, nho = 1200, 2, 3, 300
    c, X, y = data(n, p, seed=m)
    ho = folds(n, k, nho, seed=m + 1)
    ig = [2.0, 1.0]
    phis = [2.0, 6.0, 15.0]
    alphas = [0.05, 0.5]
    ta = np.array([[ph, al] + ([nu] if cov == 2 else [])
                   for ph in phis for al in alphas])
    g = ta.shape[0]
    with nngp.CvGrid(y, X, c, m, ho, ig, cov) as cv:
        r = cv.run(ta, details=True, predictions=True)
        r2 = cv.run(ta)
    # same launch shape, same sums: bitwise repeatable
    assert np.array_equal(r["scores"], r2["scores"])
    want = np.zeros((g, 2))
    for f in range(k):
        keep = np.ones(n, bool)
        keep[ho[f]] = False
        cm, Xm, ym = c[keep], X[keep], y[keep]
        c0, X0, y0 = c[ho[f]], X[ho[f]], y[ho[f]]
        nn = oracle.mkNNIndxCB(cm, m)
        nn0, _ = oracle.nn2(cm, c0, m)
        ref = oracle.cNNGP(ym, Xm.T.reshape(-1).copy(), p, len(ym),
                           cm.T.reshape(-1).copy(), ta.reshape(-1).copy(),
                           nn["nnIndx"], nn["nnIndxLU"],
                           X0.T.reshape(-1).copy(), c0.T.reshape(-1).copy(),
                           nho, nn0.T.reshape(-1).copy(), g, m, ig, cov)
        assert max_rel(r["beta.hat"][f].T, ref["beta.hat"]) < RTOL
        assert max_rel(r["ab"][f].T, ref["ab"]) < RTOL
        bref = ref["bB.inv"].T.reshape(g, p, p).transpose(0, 2, 1)
        assert max_rel(r["bB.inv"][f], bref) < RTOL
        assert max_rel(r["y.0.hat"][f].T, ref["y.0.hat"]) < RTOL
        assert max_rel(r["y.0.hat.var"][f].T, ref["y.0.hat.var"]) < RTOL
        for j in range(g):
            want[j, 0] += r_rmspe(y0, ref["y.0.hat"][:, j]) / k
            want[j, 1] += r_crps(y0, ref["y.0.hat"][:, j],
                                 ref["y.0.hat.var"][:, j]) / k
    assert max_rel(r["scores"], want) < RTOL


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(0, 64, 0), (0, 256, 0), (0, 64, 2),
                                   (0, 128, 2), (1, 128, 0)])
def test_cvgrid_every_pinned_shape_matches_oracle(oracle, nngp, shape):
    """A grid pinned to an updateBF launch shape (per lane, per wavefront)
    gives the oracle's per-fold solution and skips the autotuner."""
    print(f"cvgrid shape={shape}", flush=True)
    n, p, k, nho, m = 1500, 2, 2, 300, 10
    c, X, y = data(n, p, seed=21)
    ho = folds(n, k, nho, seed=22)
    ig = [2.0, 1.0]
    ta = np.array([[ph, al] for ph in (2.0, 9.0) for al in (0.05, 0.5)])
    g = ta.shape[0]
    with nngp.CvGrid(y, X, c, m, ho, ig, 0, shape=shape) as cv:
        r = cv.run(ta, details=True)
    assert r["tune"] == []
    assert r["bf_config"][1] == shape[1]
    for f in range(k):
        keep = np.ones(n, bool)
        keep[ho[f]] = False
        cm, Xm, ym = c[keep], X[keep], y[keep]
        c0, X0 = c[ho[f]], X[ho[f]]
        nn = oracle.mkNNIndxCB(cm, m)
        nn0, _ = oracle.nn2(cm, c0, m)
        ref = oracle.cNNGP(ym, Xm.T.reshape(-1).copy(), p, len(ym),
                           cm.T.reshape(-1).copy(), ta.reshape(-1).copy(),
                           nn["nnIndx"], nn["nnIndxLU"],
                           X0.T.reshape(-1).copy(), c0.T.reshape(-1).copy(),
                           nho, nn0.T.reshape(-1).copy(), g, m, ig, 0)
        assert max_rel(r["beta.hat"][f].T, ref["beta.hat"]) < RTOL
        assert max_rel(r["ab"][f].T, ref["ab"]) < RTOL


@pytest.mark.gpu
def test_cvgrid_autotuner_reports_candidates(nngp):
    print("cvgrid autotune", flush=True)
    n, p, k, nho, m = 3000, 3, 2, 600, 15
    c, X, y = data(n, p, seed=4)
    ho = folds(n, k, nho, seed=5)
    ta = np.array([[ph, al, 1.5] for ph in (3.0, 9.0) for al in (0.1, 1.0)])
    with nngp.CvGrid(y, X, c, m, ho, [2.0, 1.0], "matern") as cv:
        r = cv.run(ta)
    tune = r["tune"]
    kinds = {t[0] for t in tune}
    # per lane at 64/128/256 threads, staged at 64/128 (256 would need
    # 128 KiB of LDS a block), per wavefront at 1/2/4 waves
    assert kinds == {"lane", "lane+lds-z", "wave"} and len(tune) == 8
    best = min(tune, key=lambda t: t[2])
    assert r["bf_config"] == (best[0], best[1])
    assert r["graph_ms"] > 0
    assert np.all(np.isfinite(r["scores"]))
