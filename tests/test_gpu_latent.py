"""GPU parity of the latent-model pieces against the oracle
(src/nn.cpp:25-56, src/sNNGP.cpp:90-482):

  * nngp_mkUIndx is bit-exact to the oracle's scans for search types 0, 1
    and 2, on uniform and on clustered sites.
  * nngp_sNNGP with the site-order sweep follows the oracle's chain (the
    reference's sampler) under the same stream of normals and uniforms:
    beta, theta and every w sample within 1e-9, the same accepted moves.
  * nngp_sNNGP with the coloured sweep follows the oracle's chain swept in
    colour order; the colouring is rebuilt here from its rule (greedy, site
    order, lowest colour unused among the moral-graph neighbors).
  * Without a caller's generator the sweep draws Philox normals on the
    device: the same seed gives the same chain, and the chain recovers the
    simulated slope, linear predictor and w."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sites(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        cen = rng.random((20, 2))
        c = cen[rng.integers(0, 20, n)] + 0.01 * rng.standard_normal((n, 2))
    else:
        c = rng.random((n, 2))
    return c[np.argsort(c[:, 0], kind="stable")], rng


@pytest.mark.parametrize("n,m,clustered", [(50, 3, False), (2000, 10, False),
                                           (3000, 15, True), (20000, 15, False)])
def test_mkUIndx_bit_exact(oracle, n, m, clustered):
    import spnngp_amd as nn
    c, _ = _sites(n, 11, clustered)
    r = oracle.mkNNIndxCB(c, m)
    print(f"[mkUIndx] n={n} m={m} clustered={clustered}", flush=True)
    want2 = oracle.mkUIndx(n, m, r["nnIndx"], r["nnIndxLU"], 2)
    ulu = want2["u.indx.lu"]
    want1 = {k: v.copy() for k, v in want2.items()}
    for i in range(n):  # search type 1: each group reversed
        o, k = ulu[i], ulu[n + i]
        for key in ("u.indx", "ui.indx"):
            want1[key][o:o + k] = want2[key][o:o + k][::-1]
    if n <= 3000:  # the O(n^2 m) scans of search types 0 and 1
        for t, w in ((0, want2), (1, want1)):
            ref = oracle.mkUIndx(n, m, r["nnIndx"], r["nnIndxLU"], t)
            for key in w:
                assert np.array_equal(ref[key], w[key]), (t, key)
    for t, w in ((0, want2), (1, want1), (2, want2)):
        got = nn.mkUIndx(n, m, r["nnIndx"], r["nnIndxLU"], t)
        for key in w:
            assert np.array_equal(got[key], w[key]), (t, key)


def _problem(n, m, seed, cov=0, p=2):
    from oracle import oracle as O
    c, rng = _sites(n, seed)
    r = O.mkNNIndxCB(c, m)
    U = O.mkUIndx(n, m, r["nnIndx"], r["nnIndxLU"], 2)
    X = np.column_stack([np.ones(n)] + [rng.standard_normal(n)
                                        for _ in range(p - 1)])
    D = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    w = np.linalg.cholesky(1.5 * np.exp(-5.0 * D) + 1e-9 * np.eye(n)) \
        @ rng.standard_normal(n)
    beta = np.array([1.0, -2.0, 0.5][:p])
    y = X @ beta + w + np.sqrt(0.3) * rng.standard_normal(n)
    return dict(c=c, nn=r["nnIndx"], lu=r["nnIndxLU"], U=U, X=X, y=y, w=w,
                beta=beta, rng=rng, n=n, m=m, p=p, cov=cov)


def _args(P, S):
    nu = [0.5, 2.5] if P["cov"] == 2 else None
    # C carries no nugget in this model: the gaussian covariance starts at a
    # shorter range, where the neighbor systems are well conditioned enough
    # for two implementations to agree to 1e-9 over a chain
    phi0, tune = (20.0, 0.05) if P["cov"] == 3 else (5.0, 0.4)
    return (P["y"], P["X"], P["p"], P["n"], P["m"], P["c"], P["cov"],
            P["nn"], P["lu"], P["U"]["u.indx"], P["U"]["u.indx.lu"],
            P["U"]["ui.indx"], [2.0, 1.5], [2.0, 0.3], [2.0, 30.0], nu,
            np.zeros(P["p"]), 1.0, 0.5, phi0, 1.2, 0.0, 0.0, tune, 0.2, S)


def _colour_order(P):
    n, nn, lu = P["n"], P["nn"], P["lu"]
    u, ulu = P["U"]["u.indx"], P["U"]["u.indx.lu"]
    parents = [nn[lu[i]:lu[i] + lu[n + i]] for i in range(n)]
    col = np.zeros(n, int)
    for i in range(n):
        used = {col[k] for k in parents[i]}
        for j in u[ulu[i]:ulu[i] + ulu[n + i]]:
            used |= {col[k] for k in parents[j] if k < i}
        col[i] = next(c for c in range(n) if c not in used)
    return np.lexsort((np.arange(n), col)).astype(np.int32)


def _compare(got, want):
    for k in ("p.beta.samples", "p.theta.samples", "p.w.samples"):
        g, w = got[k], want[k]
        scale = np.maximum(np.abs(w), 1.0)
        assert np.max(np.abs(g - w) / scale) < 1e-9, k
    assert got["accept"] == want["accept"]


@pytest.mark.parametrize("cov,sweep", [(0, "site"), (0, "coloured"),
                                       (3, "site"), (2, "coloured"),
                                       (1, "site")])
def test_sNNGP_chain_matches_oracle(oracle, cov, sweep):
    import spnngp_amd as nn
    P = _problem(300, 8, seed=21 + cov, cov=cov)
    S = 40
    rng = np.random.default_rng(7)
    normals = rng.standard_normal(S * (P["n"] + 12))
    unifs = rng.random(S * 8)
    print(f"[sNNGP] cov={cov} sweep={sweep} n={P['n']} S={S}", flush=True)
    visit = None if sweep == "site" else _colour_order(P)
    args = _args(P, S)
    a_or = list(args)
    a_or[5] = P["c"].T.reshape(-1).copy()  # the oracle takes flat coordinates
    want = oracle.sNNGP(*a_or[:21], a_or[23], a_or[24], S,
                        oracle.TapeRng(normals, unifs), visit=visit)
    got = nn.sNNGP(*args, rng=oracle.TapeRng(normals, unifs), sweep=sweep)
    _compare(got, want)


def test_sNNGP_device_normals_are_deterministic_and_recover_truth():
    import spnngp_amd as nn
    P = _problem(600, 10, seed=41)
    S = 1200
    print(f"[sNNGP philox] n={P['n']} S={S}", flush=True)
    a = nn.sNNGP(*_args(P, S), seed=5, sweep="coloured")
    b = nn.sNNGP(*_args(P, S), seed=5, sweep="coloured")
    for k in ("p.beta.samples", "p.theta.samples", "p.w.samples"):
        assert np.array_equal(a[k], b[k])
    # the intercept trades off against the mean of w (a slow direction of
    # the Gibbs chain); the slope and the linear predictor are identified
    keep = slice(S // 3, S)
    bs = a["p.beta.samples"][:, keep]
    assert abs(bs[1].mean() - P["beta"][1]) < 4 * bs[1].std() + 0.05
    wm = a["p.w.samples"][:, keep].mean(1)
    eta = P["X"] @ bs.mean(1) + wm
    assert np.corrcoef(eta, P["X"] @ P["beta"] + P["w"])[0, 1] > 0.95
    assert np.corrcoef(wm, P["w"])[0, 1] > 0.8
    c = nn.sNNGP(*_args(P, 20), seed=6, sweep="site")
    assert np.all(np.isfinite(c["p.w.samples"]))
    assert not np.array_equal(c["p.w.samples"], a["p.w.samples"][:, :20])


def test_spNNGP_latent_front_end_follows_the_oracle(oracle):
    """spNNGP(method = "latent") (R/spNNGP.R:332-347, 409-413, 464-484):
    x-ordering, the neighbors-of-neighbors index, the sampler and w put back
    in the caller's order."""
    import spnngp_amd as nn
    n, m, S = 250, 6, 30
    rng = np.random.default_rng(17)
    c = rng.random((n, 2))  # caller's order, not x-ordered
    X = np.column_stack([np.ones(n), rng.standard_normal(n)])
    y = X @ np.array([1.0, 2.0]) + rng.standard_normal(n)
    normals = rng.standard_normal(S * (n + 12))
    unifs = rng.random(S * 8)
    print(f"[spNNGP latent] n={n} S={S}", flush=True)
    fit = nn.spNNGP(y, X, c, method="latent", n_neighbors=m,
                    starting={"beta": [0.0, 0.0], "sigma.sq": 1.0,
                              "tau.sq": 0.5, "phi": 5.0},
                    tuning={"phi": 0.3},
                    priors={"sigma.sq.IG": [2, 1], "tau.sq.IG": [2, 0.5],
                            "phi.Unif": [1, 30]},
                    n_samples=S, rng=oracle.TapeRng(normals, unifs),
                    return_neighbor_info=True, fit_rep=True,
                    sub_sample={"start": 11, "thin": 2})
    ord_ = np.argsort(c[:, 0], kind="stable")
    co = c[ord_]
    r = oracle.mkNNIndxCB(co, m)
    U = oracle.mkUIndx(n, m, r["nnIndx"], r["nnIndxLU"], 2)
    ni = fit["neighbor.info"]
    assert np.array_equal(ni["u.indx"], U["u.indx"])
    assert np.array_equal(ni["ui.indx"], U["ui.indx"])
    want = oracle.sNNGP(y[ord_], X[ord_], 2, n, m, co.T.reshape(-1).copy(), 0,
                        r["nnIndx"], r["nnIndxLU"], U["u.indx"],
                        U["u.indx.lu"], U["ui.indx"], [2, 1], [2, 0.5],
                        [1, 30], None, [0.0, 0.0], 1.0, 0.5, 5.0, 0.0, 0.3,
                        0.0, S, oracle.TapeRng(normals, unifs))
    np.testing.assert_allclose(fit["p.w.samples"][ord_], want["p.w.samples"],
                               rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(fit["p.beta.samples"], want["p.beta.samples"].T,
                               rtol=1e-9, atol=1e-9)
    assert fit["type"] == ["latent", "gaussian"]
    k = len(fit["s.indx"])
    assert fit["y.hat.samples"].shape == (n, k) == fit["y.rep.samples"].shape
    yh = fit["X"] @ fit["p.beta.samples"][fit["s.indx"] - 1].T \
        + fit["p.w.samples"][:, fit["s.indx"] - 1]
    np.testing.assert_allclose(fit["y.hat.samples"], yh, rtol=1e-12)
    f2 = nn.fitted(fit, {"start": 21, "thin": 1})
    assert f2["y.hat.samples"].shape == (n, S - 20)


@pytest.mark.parametrize("cov,m", [(0, 10), (1, 15), (3, 8), (2, 10),
                                   (0, 24)])
def test_sNNGPPredict_matches_oracle(oracle, cov, m):
    """nngp_sNNGPPredict (src/sNNGPPredict.cpp:25-241) against the oracle
    fed the same 2 q S normals: the lane kernel (m <= 16, closed-form
    covariances) and the wavefront kernel (Matern, m > 16)."""
    import spnngp_amd as nn
    rng = np.random.default_rng(100 + cov + m)
    n, q, S, p = 3000, 257, 37, 2
    c = rng.random((n, 2))
    c0 = rng.random((q, 2))
    X = np.column_stack([np.ones(n), rng.standard_normal(n)])
    X0 = np.column_stack([np.ones(q), rng.standard_normal(q)])
    d2 = ((c0[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    nn0 = np.argsort(d2, axis=1, kind="stable")[:, :m].astype(np.int32)
    beta = rng.standard_normal((p, S))
    # no nugget in C: the smooth covariances (gaussian, matern with larger
    # nu) need a short range for the neighbor systems to be conditioned
    # well enough for 1e-10 agreement between two implementations
    phi_lo, phi_hi = {3: (60.0, 120.0), 2: (20.0, 40.0)}.get(cov, (5.0, 20.0))
    theta = np.vstack([rng.uniform(0.5, 2.0, S), rng.uniform(0.1, 0.5, S),
                       rng.uniform(phi_lo, phi_hi, S)]
                      + ([rng.uniform(0.4, 1.0, S)] if cov == 2 else []))
    w = rng.standard_normal((n, S))
    z = rng.standard_normal(2 * q * S)
    print(f"[sNNGPPredict] cov={cov} m={m} q={q} S={S}", flush=True)
    got = nn.sNNGPPredict(X, None, c, n, p, m, X0, c0, q, nn0, beta, theta,
                          w, S, cov, z=z)
    y0, w0 = oracle.sNNGPPredict(X0.T.reshape(-1).copy(),
                                 c.T.reshape(-1).copy(), n, p, m,
                                 c0.T.reshape(-1).copy(), q,
                                 nn0.T.reshape(-1).copy(),
                                 beta.T.reshape(-1).copy(),
                                 theta.T.reshape(-1).copy(),
                                 w.T.reshape(-1).copy(), S, cov, z)
    np.testing.assert_allclose(got["p.w.0"], w0, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(got["p.y.0"], y0, rtol=1e-10, atol=1e-10)


def test_predict_of_a_latent_fit(oracle):
    """predict() of spNNGP(method = "latent") (R/generics.R:583-675): nn2 on
    the caller-order sites, the kept w samples, the oracle's sNNGPPredict
    on the same inputs and normals."""
    import spnngp_amd as nn
    n, m, S = 300, 8, 12
    rng = np.random.default_rng(23)
    c = rng.random((n, 2))
    X = np.column_stack([np.ones(n), rng.standard_normal(n)])
    y = X @ np.array([1.0, 2.0]) + rng.standard_normal(n)
    fit = nn.spNNGP(y, X, c, method="latent", n_neighbors=m,
                    starting={"beta": [0.0, 0.0], "sigma.sq": 1.0,
                              "tau.sq": 0.5, "phi": 5.0},
                    tuning={"phi": 0.3},
                    priors={"sigma.sq.IG": [2, 1], "tau.sq.IG": [2, 0.5],
                            "phi.Unif": [1, 30]},
                    n_samples=S, seed=3, sweep="coloured")
    q = 40
    c0 = rng.random((q, 2))
    X0 = np.column_stack([np.ones(q), rng.standard_normal(q)])
    zz = rng.standard_normal(2 * q * (S - 4))
    tape = oracle.TapeRng(zz, [])
    print(f"[predict latent] n={n} q={q}", flush=True)
    out = nn.predict(fit, X0, c0, sub_sample={"start": 5}, rng=tape)
    assert out["p.y.0"].shape == (q, S - 4) == out["p.w.0"].shape
    nn0, _ = nn.nn2(c, c0, m)
    keep = np.arange(4, S)
    y0, w0 = oracle.sNNGPPredict(
        X0.T.reshape(-1).copy(), fit["coords"].T.reshape(-1).copy(), n, 2, m,
        c0.T.reshape(-1).copy(), q, nn0.T.reshape(-1).copy(),
        fit["p.beta.samples"][keep].reshape(-1).copy(),
        fit["p.theta.samples"][keep].reshape(-1).copy(),
        fit["p.w.samples"][:, keep].T.reshape(-1).copy(), S - 4, 0, zz)
    np.testing.assert_allclose(out["p.w.0"], w0, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(out["p.y.0"], y0, rtol=1e-10, atol=1e-10)


def test_latent_edge_cases(oracle):
    """The smallest layout (n = 2, m = 1), one covariate, no samples, and a
    neighbor set that holds every predecessor (m = n - 1, the exact GP)."""
    import spnngp_amd as nn
    c = np.array([[0.1, 0.2], [0.4, 0.3]])
    r = oracle.mkNNIndxCB(c, 1)
    for t in (0, 1, 2):
        got = nn.mkUIndx(2, 1, r["nnIndx"], r["nnIndxLU"], t)
        want = oracle.mkUIndx(2, 1, r["nnIndx"], r["nnIndxLU"], t)
        for k in want:
            assert np.array_equal(got[k], want[k])
    n = 40
    P = _problem(n, n - 1, seed=8, p=1)
    out = nn.sNNGP(*_args(P, 0))
    assert out["p.w.samples"].shape == (n, 0)
    S = 15
    rng = np.random.default_rng(4)
    normals = rng.standard_normal(S * (n + 12))
    unifs = rng.random(S * 8)
    print(f"[sNNGP exact GP] n={n} m={n - 1}", flush=True)
    for sweep in ("site", "coloured"):
        visit = None if sweep == "site" else _colour_order(P)
        a_or = list(_args(P, S))
        a_or[5] = P["c"].T.reshape(-1).copy()
        want = oracle.sNNGP(*a_or[:21], a_or[23], a_or[24], S,
                            oracle.TapeRng(normals, unifs), visit=visit)
        got = nn.sNNGP(*_args(P, S), rng=oracle.TapeRng(normals, unifs),
                       sweep=sweep)
        _compare(got, want)


def test_sNNGP_verbose_prints_the_reference_report(oracle, capfd):
    import spnngp_amd as nn
    P = _problem(120, 6, seed=12)
    nn.sNNGP(*_args(P, 6), verbose=1, nReport=2, seed=3, sweep="coloured")
    out = capfd.readouterr().out
    assert "NNGP Latent model fit with 120 observations." in out
    assert "Sampled: 6 of 6, 100.00%" in out
    assert "colour classes" in out


@pytest.mark.parametrize("sig_a,tau_a", [(0.5, 8.0), (12.0, 0.6)])
def test_sNNGP_small_shapes_follow_the_oracle_gamma(oracle, sig_a, tau_a):
    """With n = 6 the Gamma shapes IG.a + n/2 are small (3.5 and 11, or 15
    and 3.6), where algorithm GD often leaves its immediate and squeeze
    tests for the quotient test and the double-exponential loop; 300
    iterations take every branch, so the library's Gamma deviates must
    follow the oracle's draw for draw."""
    import spnngp_amd as nn
    P = _problem(6, 2, seed=31)
    S = 300
    rng = np.random.default_rng(77)
    normals = rng.standard_normal(S * 40)
    unifs = rng.random(S * 40)
    a = list(_args(P, S))
    a[12], a[13] = [sig_a, 1.0], [tau_a, 0.5]
    a_or = list(a)
    a_or[5] = P["c"].T.reshape(-1).copy()
    t_or = oracle.TapeRng(normals, unifs)
    want = oracle.sNNGP(*a_or[:21], a_or[23], a_or[24], S, t_or)
    t_gpu = oracle.TapeRng(normals, unifs)
    got = nn.sNNGP(*a, rng=t_gpu)
    _compare(got, want)
    assert (t_gpu.i_n, t_gpu.i_u) == (t_or.i_n, t_or.i_u)
    # more uniforms than one per iteration: the Gamma draws went past the
    # immediate-acceptance branch
    assert t_or.i_u > 2 * S
