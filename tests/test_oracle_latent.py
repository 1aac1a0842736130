"""Pins of the oracle's latent-model pieces (src/nn.cpp:25-56,
src/util.cpp:72-185, src/sNNGP.cpp:90-482) by what the mathematics fixes:

  * mkUIndx: the three search types agree (0 and 2 ascending, 1 the same
    groups reversed), and the index is the transpose of the neighbor lists,
    rebuilt here from the (site, slot) pairs with numpy.
  * rgamma (algorithm GD): Kolmogorov-Smirnov against scipy's Gamma CDF for
    shapes in each of GD's three parameter ranges; the immediate-acceptance
    branch in closed form (normal t >= 0 gives scale (s + t/2)^2).
  * sNNGP, one iteration from a crafted stream of normals and uniforms,
    against dense linear algebra: the w sweep equals the sequence of full
    conditionals of w_i under the dense precision (I-B)^T F^-1 (I-B) +
    I/tau.sq in the visit order; beta with zero normals is the least-squares
    fit of y - w; tau.sq and sigma.sq with a zero normal are the closed
    forms scale (a - 1/2) of GD's first branch.
  * sNNGP, a long chain: the posterior recovers the simulated slope, linear
    predictor and w.
Host code only."""
import numpy as np
import pytest
from scipy import stats


def _data(n, m, seed=1, p=2):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    c = rng.random((n, 2))
    c = c[np.argsort(c[:, 0], kind="stable")]
    r = O.mkNNIndxCB(c, m)
    X = np.column_stack([np.ones(n)] + [rng.standard_normal(n)
                                        for _ in range(p - 1)])
    return c, r["nnIndx"], r["nnIndxLU"], X, rng


@pytest.mark.parametrize("n,m", [(40, 3), (300, 8), (500, 15)])
def test_mkUIndx_search_types_and_transpose(oracle, n, m):
    c, nn, lu, _, _ = _data(n, m)
    U = [oracle.mkUIndx(n, m, nn, lu, t) for t in (0, 1, 2)]
    for k in U[0]:
        assert np.array_equal(U[0][k], U[2][k])
    assert np.array_equal(U[0]["u.indx.lu"], U[1]["u.indx.lu"])
    ulu = U[0]["u.indx.lu"]
    # numpy transpose of the (site, slot) pairs
    rows = np.repeat(np.arange(n), lu[n:])
    slots = np.concatenate([np.arange(k) for k in lu[n:]])
    cols = np.concatenate([nn[lu[i]:lu[i] + lu[n + i]] for i in range(n)])
    key = np.lexsort((rows, cols))
    assert np.array_equal(U[0]["u.indx"], rows[key])
    assert np.array_equal(U[0]["ui.indx"], slots[key])
    assert np.array_equal(ulu[n:], np.bincount(cols, minlength=n))
    assert np.array_equal(ulu[:n], np.concatenate([[0], np.cumsum(ulu[n:])[:-1]]))
    for i in range(n):
        o, k = ulu[i], ulu[n + i]
        assert np.array_equal(U[1]["u.indx"][o:o + k], U[0]["u.indx"][o:o + k][::-1])
        assert np.array_equal(U[1]["ui.indx"][o:o + k], U[0]["ui.indx"][o:o + k][::-1])


@pytest.mark.parametrize("a", [1.0, 1.7, 3.6, 5.0, 12.9, 13.5, 80.0, 1500.0])
def test_rgamma_distribution(oracle, a):
    rng = np.random.default_rng(int(a * 10))
    tape = oracle.TapeRng(rng.standard_normal(60000), rng.random(200000))
    g = oracle.rgamma(tape, a, 0.7, 12000)
    assert np.all(np.isfinite(g)) and np.all(g > 0)
    assert stats.kstest(g, stats.gamma(a, scale=0.7).cdf).pvalue > 1e-4


def test_rgamma_immediate_acceptance_closed_form(oracle):
    for a, t in [(2.0, 0.0), (10.0, 0.3), (500.0, 1.7)]:
        tape = oracle.TapeRng([t], [])
        s = np.sqrt(a - 0.5)
        assert oracle.rgamma(tape, a, 3.0)[0] == pytest.approx(
            3.0 * (s + 0.5 * t) ** 2, rel=1e-14)
        assert tape.i_n == 1 and tape.i_u == 0


def _dense_Q(B, F, nn, lu, n):
    IB = np.eye(n)
    for i in range(n):
        for k in range(lu[n + i]):
            IB[i, nn[lu[i] + k]] -= B[lu[i] + k]
    return IB.T @ np.diag(1.0 / F) @ IB


@pytest.mark.parametrize("cov,visit_kind", [(0, "site"), (0, "perm"),
                                            (3, "site"), (2, "perm")])
def test_sNNGP_one_iteration_against_dense_algebra(oracle, cov, visit_kind):
    n, m, p = 60, 5, 2
    c, nn, lu, X, rng = _data(n, m, seed=3, p=p)
    U = oracle.mkUIndx(n, m, nn, lu, 2)
    y = X @ np.array([1.0, -2.0]) + rng.standard_normal(n)
    th0 = dict(s2=2.0, t2=0.5, phi=4.0, nu=1.2)
    beta0 = np.array([0.7, -1.5])
    zw = rng.standard_normal(n)
    # stream: n sweep normals, p beta normals (0), one normal for each Gamma
    # (0: immediate acceptance), the phi (and nu) proposal normal (0), and
    # the uniform of the Metropolis test
    normals = np.concatenate([zw, np.zeros(p + 2 + (2 if cov == 2 else 1))])
    tape = oracle.TapeRng(normals, [0.5])
    visit = None if visit_kind == "site" else rng.permutation(n).astype(np.int32)
    sIG, tIG = np.array([2.0, 1.5]), np.array([2.5, 0.4])
    phiU, nuU = np.array([1.0, 30.0]), np.array([0.5, 2.5])
    out = oracle.sNNGP(
        y, X, p, n, m, c.T.reshape(-1).copy(), cov, nn, lu, U["u.indx"],
        U["u.indx.lu"], U["ui.indx"], sIG, tIG, phiU, nuU, beta0,
        th0["s2"], th0["t2"], th0["phi"], th0["nu"], 0.1, 0.1, 1, tape,
        visit=visit)
    theta = [th0["s2"], 0.0, th0["phi"]] + ([th0["nu"]] if cov == 2 else [])
    B = np.zeros(oracle.n_indx(n, m))
    F = np.zeros(n)
    ld = np.zeros(1)
    import ctypes as C
    d = C.c_double()
    oracle._check(oracle.lib().or_updateBF(
        B, F, c.T.reshape(-1).copy(), nn, lu, n, m,
        np.array(theta + [0.0] * (4 - len(theta))), cov, 1, C.byref(d)))
    Q = _dense_Q(B, F, nn, lu, n)
    # the sweep: full conditionals in the visit order
    w = np.zeros(n)
    order = np.arange(n) if visit is None else visit
    for i in order:
        prec = 1.0 / th0["t2"] + Q[i, i]
        rhs = (y[i] - X[i] @ beta0) / th0["t2"] - (Q[i] @ w - Q[i, i] * w[i])
        w[i] = rhs / prec + zw[i] / np.sqrt(prec)
    np.testing.assert_allclose(out["p.w.samples"][:, 0], w, rtol=1e-11,
                               atol=1e-12)
    beta1 = np.linalg.lstsq(X, y - w, rcond=None)[0]
    np.testing.assert_allclose(out["p.beta.samples"][:, 0], beta1, rtol=1e-11)
    ss = np.sum((y - w - X @ beta1) ** 2)
    tau1 = (tIG[1] + 0.5 * ss) / (tIG[0] + n / 2.0 - 0.5)
    a = w @ Q @ w
    sig1 = (sIG[1] + 0.5 * a * th0["s2"]) / (sIG[0] + n / 2.0 - 0.5)
    th = out["p.theta.samples"][:, 0]
    assert th[1] == pytest.approx(tau1, rel=1e-11)
    assert th[0] == pytest.approx(sig1, rel=1e-11)
    assert th[2] == pytest.approx(th0["phi"], rel=1e-12)  # zero step
    assert out["accept"] == 1
    assert tape.i_n == normals.size and tape.i_u == 1


def test_sNNGP_posterior_recovers_truth(oracle):
    n, m, p = 400, 10, 2
    c, nn, lu, X, rng = _data(n, m, seed=5, p=p)
    U = oracle.mkUIndx(n, m, nn, lu, 2)
    D = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    w_true = np.linalg.cholesky(2.0 * np.exp(-6.0 * D) + 1e-10 * np.eye(n)) \
        @ rng.standard_normal(n)
    beta = np.array([1.0, 3.0])
    y = X @ beta + w_true + np.sqrt(0.2) * rng.standard_normal(n)
    S = 1500
    tape = oracle.TapeRng(rng.standard_normal(S * (n + 12)),
                          rng.random(S * 8))
    out = oracle.sNNGP(
        y, X, p, n, m, c.T.reshape(-1).copy(), 0, nn, lu, U["u.indx"],
        U["u.indx.lu"], U["ui.indx"], [2.0, 2.0], [2.0, 0.2], [3.0, 30.0],
        None, np.zeros(p), 1.0, 1.0, 6.0, 0.0, 0.3, 0.0, S, tape)
    # the intercept trades off against the mean of w; the slope and the
    # linear predictor are identified
    keep = slice(S // 3, S)
    bs = out["p.beta.samples"][:, keep]
    assert abs(bs[1].mean() - beta[1]) < 4 * bs[1].std() + 0.05
    wm = out["p.w.samples"][:, keep].mean(1)
    assert np.corrcoef(X @ bs.mean(1) + wm, X @ beta + w_true)[0, 1] > 0.95
    assert np.corrcoef(wm, w_true)[0, 1] > 0.8
    assert 0 < out["accept"] < S


@pytest.mark.parametrize("cov", [0, 2, 3])
def test_sNNGPPredict_is_the_dense_conditional_at_m_eq_n(oracle, cov):
    """With every fitted site a neighbor, w0 = the GP conditional mean of w
    at the new site given w at the fitted sites (z = 0), its sd is
    sqrt(s2 - c^T C^-1 c) (z = 1), and y0 adds x0.beta + sqrt(t2) z2."""
    rng = np.random.default_rng(9 + cov)
    n, q, S, p = 12, 5, 3, 2
    c = rng.random((n, 2))
    c0 = rng.random((q, 2))
    X0 = np.column_stack([np.ones(q), rng.standard_normal(q)])
    nn0 = np.array([np.argsort(((c - c0[i]) ** 2).sum(1)) for i in range(q)],
                   np.int32)  # q x n, all sites
    beta = rng.standard_normal((p, S))
    theta = np.vstack([[1.5, 2.0, 0.8], [0.3, 0.5, 0.2], [3.0, 5.0, 8.0]]
                      + ([[0.7, 1.3, 2.1]] if cov == 2 else []))
    w = rng.standard_normal((n, S))
    from scipy import special
    args = (X0.T.reshape(-1).copy(), c.T.reshape(-1).copy(), n, p, n,
            c0.T.reshape(-1).copy(), q, nn0.T.reshape(-1).copy(),
            beta.T.reshape(-1).copy(), theta.T.reshape(-1).copy(),
            w.T.reshape(-1).copy(), S, cov)
    z0 = np.zeros(2 * q * S)
    y0a, w0a = oracle.sNNGPPredict(*args, z0)
    z1 = np.concatenate([np.ones(q * S), np.full(q * S, 2.0)])
    y0b, w0b = oracle.sNNGPPredict(*args, z1)

    def rho(D, th):
        phi = th[2]
        if cov == 0:
            return np.exp(-phi * D)
        if cov == 3:
            return np.exp(-(phi * D) ** 2)
        nu = th[3]
        x = np.where(D > 0, D * phi, 1.0)
        r = x ** nu * special.kv(nu, x) / (2 ** (nu - 1) * special.gamma(nu))
        return np.where(D > 0, r, 1.0)

    D = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    for s in range(S):
        th = theta[:, s]
        Cd = th[0] * rho(D, th)
        for i in range(q):
            c_i = th[0] * rho(np.sqrt(((c - c0[i]) ** 2).sum(1)), th)
            u = np.linalg.solve(Cd, c_i)
            mean = u @ w[:, s]
            sd = np.sqrt(th[0] - u @ c_i)
            assert w0a[i, s] == pytest.approx(mean, rel=1e-8, abs=1e-9)
            assert w0b[i, s] - w0a[i, s] == pytest.approx(sd, rel=1e-6)
            xb = X0[i] @ beta[:, s]
            assert y0a[i, s] == pytest.approx(xb + mean, rel=1e-8, abs=1e-9)
            assert y0b[i, s] - w0b[i, s] - xb == pytest.approx(
                2.0 * np.sqrt(th[1]), rel=1e-10)
