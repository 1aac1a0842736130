"""Pins the oracle's drivers (rNNGP, rNNGPPredict, cNNGP) and the R-side
scoring with independent computations: dense Gaussian algebra in numpy,
scipy's inverse-gamma density, and numerical integration."""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.linalg as sla
import scipy.stats as ss

from conftest import uniform_coords


def dense_cov(c, s2, t2, phi):
    D = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    return s2 * np.exp(-phi * D) + t2 * np.eye(len(c))


def nngp_precision(B, F, nnIndx, lu, n):
    IB = np.eye(n)
    for i in range(n):
        for j in range(lu[n + i]):
            IB[i, nnIndx[lu[i] + j]] -= B[lu[i] + j]
    return IB.T @ np.diag(1.0 / F) @ IB


def log_target(theta, X, y, beta, P, logdet, priors):
    """Unnormalised log posterior on the sampler's scale (log sigma.sq,
    log tau.sq, logit phi): Gaussian likelihood, inverse-gamma and uniform
    priors from scipy, plus the log-Jacobians of the transforms."""
    s2, t2, phi = theta
    r = y - X @ beta
    ll = -0.5 * logdet - 0.5 * r @ P @ r
    (sa, sb), (ta, tb), (pa, pb) = priors
    lp = (ss.invgamma.logpdf(s2, sa, scale=sb) + np.log(s2)
          + ss.invgamma.logpdf(t2, ta, scale=tb) + np.log(t2)
          + ss.uniform.logpdf(phi, pa, pb - pa)
          + np.log((phi - pa) * (pb - phi) / (pb - pa)))
    return ll + lp


@pytest.mark.parametrize("m", [5, 29])
def test_rNNGP_first_iteration_by_dense_algebra(oracle, m):
    n, p = 30, 2
    rng = np.random.default_rng(4)
    c = uniform_coords(n, 4)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, 2.0] + rng.normal(size=n)
    nn = oracle.mkNNIndx(c, m)
    L, lu = nn["nnIndx"], nn["nnIndxLU"]
    flat = c.T.reshape(-1).copy()
    th0 = np.array([1.5, 0.4, 5.0])
    tun = np.array([0.3, 0.2, 0.4])
    priors = [(2.0, 1.0), (2.0, 0.5), (1.0, 20.0)]
    z = rng.normal(size=p + 3)
    for u in (1e-9, 0.999999):  # one accepted, one rejected proposal
        out = oracle.rNNGP(y, X, p, n, m, flat, 0, L, lu,
                           sigmaSqIG=list(priors[0]), tauSqIG=list(priors[1]),
                           phiUnif=list(priors[2]), nuUnif=None,
                           betaStarting=np.zeros(p), sigmaSqStarting=th0[0],
                           tauSqStarting=th0[1], phiStarting=th0[2],
                           nuStarting=0.0, sigmaSqTuning=tun[0],
                           tauSqTuning=tun[1], phiTuning=tun[2], nuTuning=0.0,
                           nSamples=1, rng=oracle.TapeRng(z, [u]))
        # beta | theta0 ~ N(G^-1 h, G^-1), G = X'PX, h = X'Py
        B0, F0, ld0 = oracle.updateBF(flat, L, lu, n, m, list(th0), 0)
        P0 = nngp_precision(B0, F0, L, lu, n)
        G = X.T @ P0 @ X
        V = np.linalg.inv(G)
        beta = V @ (X.T @ P0 @ y) + np.linalg.cholesky(V) @ z[:p]
        np.testing.assert_allclose(out["p.beta.samples"][:, 0], beta,
                                   rtol=1e-10)
        # proposal in draw order phi, sigma.sq, tau.sq
        pa, pb = priors[2]
        lg = np.log((th0[2] - pa) / (pb - th0[2])) + tun[2] * z[p]
        cand = np.array([np.exp(np.log(th0[0]) + tun[0] * z[p + 1]),
                         np.exp(np.log(th0[1]) + tun[1] * z[p + 2]),
                         pb - (pb - pa) / (1 + np.exp(lg))])
        B1, F1, ld1 = oracle.updateBF(flat, L, lu, n, m, list(cand), 0)
        P1 = nngp_precision(B1, F1, L, lu, n)
        delta = (log_target(cand, X, y, beta, P1, ld1, priors)
                 - log_target(th0, X, y, beta, P0, ld0, priors))
        expect = cand if u <= np.exp(delta) else th0
        np.testing.assert_allclose(out["p.theta.samples"][:, 0], expect,
                                   rtol=1e-12)
        assert out["accept"] == int(u <= np.exp(delta))


def test_rNNGP_exact_gp_likelihood(oracle):
    """m = n-1: the log-det and quadratic form the sampler uses are those of
    the dense Gaussian N(X beta, s2 R + t2 I)."""
    n = 25
    c = uniform_coords(n, 9)
    nn = oracle.mkNNIndx(c, n - 1)
    th = [2.0, 0.3, 4.0]
    B, F, ld = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                               nn["nnIndxLU"], n, n - 1, th, 0)
    C = dense_cov(c, *th)
    r = np.random.default_rng(1).normal(size=n)
    q = oracle.Q(B, F, r, r, nn["nnIndx"], nn["nnIndxLU"])
    ref = ss.multivariate_normal(np.zeros(n), C).logpdf(r)
    assert -0.5 * ld - 0.5 * q - 0.5 * n * np.log(2 * np.pi) == \
        pytest.approx(ref, rel=1e-11)


def test_rNNGP_draw_counts(oracle):
    n, m, p, S = 60, 5, 2, 40
    rng = np.random.default_rng(2)
    c = uniform_coords(n, 2)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, 1.0] + rng.normal(size=n)
    nn = oracle.mkNNIndx(c, m)
    tape = oracle.TapeRng(rng.normal(size=1000), rng.uniform(size=1000))
    oracle.rNNGP(y, X, p, n, m, c.T.reshape(-1).copy(), 0, nn["nnIndx"],
                 nn["nnIndxLU"], sigmaSqIG=[2, 1], tauSqIG=[2, 1],
                 phiUnif=[1, 30], nuUnif=None, betaStarting=np.zeros(p),
                 sigmaSqStarting=1.0, tauSqStarting=1.0, phiStarting=5.0,
                 nuStarting=0.0, sigmaSqTuning=0.1, tauSqTuning=0.1,
                 phiTuning=0.1, nuTuning=0.0, nSamples=S, rng=tape)
    # p beta normals + 3 proposal normals and one uniform per iteration
    assert tape.i_n == S * (p + 3)
    assert tape.i_u == S
    # a zero tuning fixes the parameter and draws nothing for it
    tape = oracle.TapeRng(rng.normal(size=1000), rng.uniform(size=1000))
    out = oracle.rNNGP(y, X, p, n, m, c.T.reshape(-1).copy(), 0,
                       nn["nnIndx"], nn["nnIndxLU"], sigmaSqIG=[2, 1],
                       tauSqIG=[2, 1], phiUnif=[1, 30], nuUnif=None,
                       betaStarting=np.zeros(p), sigmaSqStarting=1.0,
                       tauSqStarting=1.0, phiStarting=5.0, nuStarting=0.0,
                       sigmaSqTuning=0.1, tauSqTuning=0.0, phiTuning=0.1,
                       nuTuning=0.0, nSamples=S, rng=tape)
    assert tape.i_n == S * (p + 2)
    assert np.all(out["p.theta.samples"][1] == 1.0)


def test_rNNGP_recovers_parameters(oracle):
    n, m, S = 300, 10, 400
    rng = np.random.default_rng(6)
    c = uniform_coords(n, 6)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    C = dense_cov(c, 5.0, 0.0, 6.0) + 1e-10 * np.eye(n)
    y = X @ [1.0, 5.0] + np.linalg.cholesky(C) @ rng.normal(size=n) \
        + rng.normal(size=n)
    nn = oracle.mkNNIndxCB(c, m)
    tape = oracle.TapeRng(rng.normal(size=S * 6), rng.uniform(size=S))
    out = oracle.rNNGP(y, X, 2, n, m, c.T.reshape(-1).copy(), 0,
                       nn["nnIndx"], nn["nnIndxLU"], sigmaSqIG=[2, 5],
                       tauSqIG=[2, 1], phiUnif=[3, 30], nuUnif=None,
                       betaStarting=np.linalg.lstsq(X, y, rcond=None)[0],
                       sigmaSqStarting=5.0, tauSqStarting=1.0,
                       phiStarting=6.0, nuStarting=0.0, sigmaSqTuning=0.3,
                       tauSqTuning=0.3, phiTuning=0.3, nuTuning=0.0,
                       nSamples=S, rng=tape, nThreads=4)
    b = out["p.beta.samples"][:, S // 2:].mean(1)
    assert abs(b[1] - 5.0) < 0.3
    assert 0.05 < out["accept"] / S < 0.95


def test_rNNGPPredict_is_dense_kriging_when_m_is_n(oracle):
    n, p, q, S = 40, 2, 7, 3
    rng = np.random.default_rng(3)
    c = rng.uniform(size=(n, 2))
    c0 = rng.uniform(size=(q, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, -1.0] + rng.normal(size=n)
    X0 = np.column_stack([np.ones(q), rng.normal(size=q)])
    beta = rng.normal(size=(p, S))
    th = np.array([[1.5, 2.0, 0.7], [0.3, 0.1, 0.5], [4.0, 9.0, 2.0]])
    nn0 = np.tile(np.arange(n), (q, 1)).astype(np.int32)
    args = (X.T.reshape(-1).copy(), y, c.T.reshape(-1).copy(), n, p, n,
            X0.T.reshape(-1).copy(), c0.T.reshape(-1).copy(), q,
            nn0.T.reshape(-1).copy(), beta.T.reshape(-1).copy(),
            th.T.reshape(-1).copy(), S, 0)
    mean = oracle.rNNGPPredict(*args, np.zeros(q * S))
    plus = oracle.rNNGPPredict(*args, np.ones(q * S))
    for s in range(S):
        s2, t2, phi = th[:, s]
        C = dense_cov(c, s2, t2, phi)
        c0v = s2 * np.exp(-phi * np.sqrt(((c0[:, None] - c[None]) ** 2)
                                         .sum(-1)))
        w = sla.solve(C, c0v.T, assume_a="pos")
        m_ref = X0 @ beta[:, s] + w.T @ (y - X @ beta[:, s])
        sd_ref = np.sqrt(s2 + t2 - np.einsum("ij,ji->i", c0v, w))
        np.testing.assert_allclose(mean[:, s], m_ref, rtol=1e-10)
        np.testing.assert_allclose(plus[:, s] - mean[:, s], sd_ref,
                                   rtol=1e-9)


def test_cNNGP_is_dense_gls_when_m_is_n_minus_1(oracle):
    n, p, n0 = 35, 2, 6
    rng = np.random.default_rng(8)
    c = uniform_coords(n, 8)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [2.0, 1.0] + rng.normal(size=n)
    c0 = rng.uniform(size=(n0, 2))
    X0 = np.column_stack([np.ones(n0), rng.normal(size=n0)])
    m = n - 1
    nn = oracle.mkNNIndx(c, m)
    nn0, _ = oracle.nn2(c, c0, m)
    phi, alpha, ig = 3.0, 0.4, [2.0, 1.5]
    out = oracle.cNNGP(y, X.T.reshape(-1).copy(), p, n,
                       c.T.reshape(-1).copy(), np.array([phi, alpha]),
                       nn["nnIndx"], nn["nnIndxLU"], X0.T.reshape(-1).copy(),
                       c0.T.reshape(-1).copy(), n0, nn0.T.reshape(-1).copy(),
                       1, m, ig, 0)
    V = dense_cov(c, 1.0, alpha, phi)
    Vi = np.linalg.inv(V)
    G = X.T @ Vi @ X
    bhat = np.linalg.solve(G, X.T @ Vi @ y)
    a = ig[0] + n / 2
    b = ig[1] + 0.5 * (y @ Vi @ y - bhat @ X.T @ Vi @ y)
    np.testing.assert_allclose(out["beta.hat"][:, 0], bhat, rtol=1e-10)
    assert out["ab"][0, 0] == pytest.approx(a, rel=1e-15)
    assert out["ab"][1, 0] == pytest.approx(b, rel=1e-10)
    np.testing.assert_allclose(out["bB.inv"][:, 0].reshape(p, p).T,
                               np.linalg.inv(G), rtol=1e-10)
    for i in range(n0):
        N = nn0[i]
        CN = dense_cov(c[N], 1.0, alpha, phi)
        c0v = np.exp(-phi * np.sqrt(((c[N] - c0[i]) ** 2).sum(-1)))
        w = np.linalg.solve(CN, c0v)
        yh = X0[i] @ bhat + w @ (y[N] - X[N] @ bhat)
        u = X0[i] - X[N].T @ w
        var = b / (a - 1) * (u @ np.linalg.solve(G, u) + 1 + alpha - w @ c0v)
        assert out["y.0.hat"][i, 0] == pytest.approx(yh, rel=1e-10)
        assert out["y.0.hat.var"][i, 0] == pytest.approx(var, rel=1e-10)


@pytest.mark.parametrize("cov,theta", [(0, [2.0, 0.3, 5.0]),
                                       (2, [1.0, 0.5, 4.0, 1.3])])
def test_rNNGPReplicated_is_the_exact_draw_when_m_is_n_minus_1(oracle, cov,
                                                               theta):
    """At m = n - 1 the NNGP is the exact GP: the replicate is
    v = (I - B)^-1 F^1/2 z + X beta, and (I - B)^-1 F (I - B)^-T equals the
    dense covariance s2 R + t2 I.  A dropped nugget, a wrong sign or a
    transposed index breaks one of the two."""
    n, p, S = 30, 2, 3
    m = n - 1
    c = uniform_coords(n, 8)
    rng = np.random.default_rng(8)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    beta = rng.normal(size=(S, p))
    th = np.array([theta] * S, float)
    th[:, 2] *= np.array([1.0, 1.3, 0.8])
    nn = oracle.mkNNIndx(c, m)
    z = rng.normal(size=n * S)
    y = oracle.rNNGPReplicated(X.T.reshape(-1).copy(), p, n, m,
                               c.T.reshape(-1).copy(), cov, nn["nnIndx"],
                               nn["nnIndxLU"], beta.T.reshape(-1).copy(),
                               th.reshape(-1).copy(), S,
                               oracle.TapeRng(z, []))
    D = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    for s in range(S):
        B, F, _ = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, th[s], cov)
        IB = np.eye(n)
        lu = nn["nnIndxLU"]
        for i in range(n):
            for j in range(lu[n + i]):
                IB[i, nn["nnIndx"][lu[i] + j]] -= B[lu[i] + j]
        W = np.linalg.solve(IB, np.diag(np.sqrt(F)))
        R = np.vectorize(lambda d: oracle.spCor(d, th[s, 2],
                                                th[s, 3] if cov == 2 else 0.0,
                                                cov))(D)
        Cd = th[s, 0] * R + th[s, 1] * np.eye(n)
        assert np.allclose(W @ W.T, Cd, rtol=1e-9, atol=1e-11)
        want = W @ z[s * n:(s + 1) * n] + X @ beta[s]
        assert np.allclose(y[:, s], want, rtol=1e-10, atol=1e-12)
