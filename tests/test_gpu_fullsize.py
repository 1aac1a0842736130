"""Parity at the headline size (n = 1e6 sites, m = 15, exponential,
x-ordered uniform sites), where the oracle still finishes in seconds with
OpenMP:

  * mkNNIndxCB: neighbor lists, offsets and distances bit for bit;
  * updateBF: B, F and log det within 1e-10 of the oracle;
  * the model pass's cross-products: Q(r, r) against the oracle's Q on the
    same B, F, and Q's bilinearity in its first argument;
  * mkUIndx: bit for bit;
  * rNNGPPredict: 500 new sites x 8 samples within 1e-10.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, M = 1_000_000, 15
THETA = [2.0, 0.2, 12.0]
NT = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def big(oracle):
    import spnngp_amd as nn
    rng = np.random.default_rng(2)
    c = rng.uniform(size=(N, 2))
    c = c[np.argsort(c[:, 0], kind="stable")]
    X = np.column_stack([np.ones(N), rng.normal(size=(N, 2))])
    y = X @ np.array([1.0, 2.0, -1.0]) + rng.normal(size=N)
    print(f"[fullsize] n={N} m={M}: device kNN", flush=True)
    idx = nn.mkNNIndxCB(c, M)
    return dict(c=c, X=X, y=y, idx=idx, cflat=c.T.reshape(-1).copy(), rng=rng)


def test_knn_bit_exact_at_full_size(oracle, big):
    print("[fullsize] oracle kNN", flush=True)
    ref = oracle.mkNNIndxCB(big["c"], M, NT)
    for k in ("nnIndx", "nnIndxLU", "nnDist"):
        assert np.array_equal(big["idx"][k], ref[k]), k


def test_updateBF_at_full_size(oracle, big):
    import ctypes as C
    import spnngp_amd as nn
    idx = big["idx"]
    print("[fullsize] updateBF", flush=True)
    B, F, ld = nn.updateBF(big["c"], idx["nnIndx"], idx["nnIndxLU"], M, THETA,
                           "exponential")
    Bo = np.zeros(oracle.n_indx(N, M))
    Fo = np.zeros(N)
    d = C.c_double()
    oracle._check(oracle.lib().or_updateBF(
        Bo, Fo, big["cflat"], idx["nnIndx"], idx["nnIndxLU"], N, M,
        np.array(THETA + [0.0]), 0, NT, C.byref(d)))
    assert np.max(np.abs(F - Fo) / np.abs(Fo)) < 1e-10
    assert np.max(np.abs(B - Bo)) < 1e-10 * np.max(np.abs(Bo))
    assert ld == pytest.approx(d.value, rel=1e-10)


def test_pass_cross_products_at_full_size(oracle, big):
    import spnngp_amd as nn
    idx, X, y = big["idx"], big["X"], big["y"]
    beta = np.array([0.9, 2.1, -1.05])
    print("[fullsize] model pass", flush=True)
    with nn.Model(big["c"], idx["nnIndx"], idx["nnIndxLU"], M, "exponential",
                  X, y) as md:
        out = md.update(THETA, slot=0, beta=beta)
        q = out["qResid"]
        B, F = md.get_bf(0)
    r = X @ beta - y
    qo = oracle.Q(B, F, r, r, idx["nnIndx"], idx["nnIndxLU"])
    assert q == pytest.approx(qo, rel=1e-10)
    # bilinear in the first argument: Q(u + 2v, w) = Q(u, w) + 2 Q(v, w)
    rng = big["rng"]
    u, v, w = rng.normal(size=(3, N))
    qq = [nn.Q(B, F, a, w, idx["nnIndx"], idx["nnIndxLU"], M)
          for a in (u + 2 * v, u, v)]
    assert qq[0] == pytest.approx(qq[1] + 2 * qq[2], rel=1e-9, abs=1e-6)


def test_mkUIndx_at_full_size(oracle, big):
    import spnngp_amd as nn
    idx = big["idx"]
    print("[fullsize] mkUIndx", flush=True)
    got = nn.mkUIndx(N, M, idx["nnIndx"], idx["nnIndxLU"], 2)
    ref = oracle.mkUIndx(N, M, idx["nnIndx"], idx["nnIndxLU"], 2)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


def test_rNNGPPredict_at_full_size(oracle, big):
    import spnngp_amd as nn
    rng = big["rng"]
    q, S = 500, 8
    c0 = rng.uniform(size=(q, 2))
    X0 = np.column_stack([np.ones(q), rng.normal(size=(q, 2))])
    nn0, _ = nn.nn2(big["c"], c0, M)
    beta = np.tile(np.array([[1.0], [2.0], [-1.0]]), (1, S)) \
        + 0.01 * rng.normal(size=(3, S))
    theta = np.tile(np.array([THETA]).T, (1, S)) \
        * np.exp(0.02 * rng.normal(size=(3, S)))
    z = rng.normal(size=q * S)
    print("[fullsize] kriging", flush=True)
    got = nn.rNNGPPredict(big["X"], big["y"], big["c"], N, 3, M, X0, c0, q,
                          nn0, beta, theta, S, "exponential", z=z)["p.y.0"]
    ref = oracle.rNNGPPredict(big["X"].T.reshape(-1).copy(), big["y"],
                              big["cflat"], N, 3, M, X0.T.reshape(-1).copy(),
                              c0.T.reshape(-1).copy(), q,
                              nn0.T.reshape(-1).copy(),
                              beta.T.reshape(-1).copy(),
                              theta.T.reshape(-1).copy(), S, 0, z,
                              nThreads=NT)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10)
