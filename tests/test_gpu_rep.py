"""rNNGPReplicated (src/rNNGPReplicated.cpp:78-230) on the device against the
oracle's restatement, fed the same normals: through a caller's generator
(the reference's order: all n sites of a sample, then the next sample) and
through the library's counter-based stream."""
import numpy as np
import pytest

from conftest import uniform_coords

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def problem(n, m, p, S, cov, seed):
    rng = np.random.default_rng(seed)
    c = uniform_coords(n, seed)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))])
    beta = rng.normal(size=(S, p))
    nTheta = 4 if cov == 2 else 3
    th = np.vstack([np.exp(rng.normal(np.log(1.5), 0.2, S)),
                    np.exp(rng.normal(np.log(0.3), 0.2, S)),
                    np.exp(rng.normal(np.log(8.0), 0.2, S))]
                   + ([rng.uniform(0.6, 2.2, S)] if cov == 2 else []))
    assert th.shape == (nTheta, S)
    return c, X, beta, th


def oracle_rep(oracle, c, X, beta, th, nn, m, cov, z):
    n, p = X.shape
    S = beta.shape[0]
    return oracle.rNNGPReplicated(
        X.T.reshape(-1).copy(), p, n, m, c.T.reshape(-1).copy(), cov,
        nn["nnIndx"], nn["nnIndxLU"], beta.T.reshape(-1).copy(),
        th.T.reshape(-1).copy(), S, oracle.TapeRng(z, []))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("cov,m", [(0, 10), (2, 12), (1, 20), (3, 5)])
def test_replicates_follow_the_callers_stream(oracle, nngp, cov, m):
    """Per-lane (m <= 16, closed forms) and wavefront (m > 16, Bessel
    matern) updateBF both feed the level-scheduled draw."""
    n, p, S = 3000, 2, 5
    print(f"rep cov={cov} m={m}", flush=True)
    c, X, beta, th = problem(n, m, p, S, cov, 30 + m)
    nn = oracle.mkNNIndxCB(c, m)
    z = np.random.default_rng(1).normal(size=n * S)
    want = oracle_rep(oracle, c, X, beta, th, nn, m, cov, z)
    got = nngp.rNNGPReplicated(X, p, n, m, c, cov, nn["nnIndx"],
                               nn["nnIndxLU"], beta, th, S,
                               rng=oracle.TapeRng(z, []))["y.rep.samples"]
    assert got.shape == (n, S)
    assert rel(got, want) < RTOL


def test_replicates_follow_the_callers_stream_in_two_groups(oracle, nngp):
    """300 samples with the caller's generator: two launch groups through
    one output buffer, each copied out before the next starts, with the
    normals drawn sample after sample across the group boundary."""
    n, m, p, S = 300, 6, 2, 300
    print("rep caller's stream, two groups", flush=True)
    c, X, beta, th = problem(n, m, p, S, 0, 41)
    nn = oracle.mkNNIndxCB(c, m)
    z = np.random.default_rng(2).normal(size=n * S)
    want = oracle_rep(oracle, c, X, beta, th, nn, m, 0, z)
    got = nngp.rNNGPReplicated(X, p, n, m, c, 0, nn["nnIndx"],
                               nn["nnIndxLU"], beta, th, S,
                               rng=oracle.TapeRng(z, []))["y.rep.samples"]
    assert got.shape == (n, S)
    for s in (0, 255, 256, 299):
        assert rel(got[:, s], want[:, s]) < RTOL


def test_replicates_with_the_counter_based_stream(oracle, nngp):
    n, m, p, S, seed = 2500, 8, 3, 4, 77
    c, X, beta, th = problem(n, m, p, S, 0, 5)
    nn = oracle.mkNNIndxCB(c, m)
    z = np.array([oracle.lib().or_philox_normal(seed, i, s)
                  for s in range(S) for i in range(n)])
    want = oracle_rep(oracle, c, X, beta, th, nn, m, 0, z)
    got = nngp.rNNGPReplicated(X, p, n, m, c, 0, nn["nnIndx"],
                               nn["nnIndxLU"], beta, th, S,
                               seed=seed)["y.rep.samples"]
    assert rel(got, want) < RTOL


def test_replicates_deep_dag_many_samples(oracle, nngp):
    """n = 2e4 (about 1,200 dependency levels); the 40 samples run as 40
    concurrent blocks of one launch."""
    n, m, p, S = 20000, 15, 2, 40
    c, X, beta, th = problem(n, m, p, S, 0, 9)
    nn = oracle.mkNNIndxCB(c, m)
    z = np.random.default_rng(2).normal(size=n * S)
    got = nngp.rNNGPReplicated(X, p, n, m, c, 0, nn["nnIndx"],
                               nn["nnIndxLU"], beta, th, S,
                               rng=oracle.TapeRng(z, []))["y.rep.samples"]
    for s in (0, 17, 39):  # the oracle on three of the samples
        want = oracle_rep(oracle, c, X, beta[s:s + 1], th[:, s:s + 1], nn, m,
                          0, z[s * n:(s + 1) * n])
        assert rel(got[:, s], want[:, 0]) < RTOL


def test_replicates_in_several_groups(oracle, nngp):
    """300 samples exceed one launch's 256 blocks: the second group starts at
    sample 256 (its beta rows, theta columns and stream keys)."""
    n, m, p, S, seed = 600, 6, 2, 300, 11
    c, X, beta, th = problem(n, m, p, S, 0, 12)
    nn = oracle.mkNNIndxCB(c, m)
    got = nngp.rNNGPReplicated(X, p, n, m, c, 0, nn["nnIndx"],
                               nn["nnIndxLU"], beta, th, S,
                               seed=seed)["y.rep.samples"]
    for s in (0, 255, 256, 299):
        z = np.array([oracle.lib().or_philox_normal(seed, i, s)
                      for i in range(n)])
        want = oracle_rep(oracle, c, X, beta[s:s + 1], th[:, s:s + 1], nn, m,
                          0, z)
        assert rel(got[:, s], want[:, 0]) < RTOL


@pytest.mark.gpu
def test_replicates_in_three_groups_reuse_the_output_buffers(oracle, nngp):
    """600 samples in groups of 256, 256 and 88: each group's replicates
    leave while the next one computes, from two device buffers used in
    turn, so the third group writes the first group's buffer after its
    copy.  Samples of every group, at both ends, against the oracle."""
    n, m, p, S, seed = 300, 5, 2, 600, 13
    print("replicates three groups", flush=True)
    c, X, beta, th = problem(n, m, p, S, 0, 14)
    nn = oracle.mkNNIndxCB(c, m)
    got = nngp.rNNGPReplicated(X, p, n, m, c, 0, nn["nnIndx"],
                               nn["nnIndxLU"], beta, th, S,
                               seed=seed)["y.rep.samples"]
    for s in (0, 255, 256, 511, 512, 599):
        z = np.array([oracle.lib().or_philox_normal(seed, i, s)
                      for i in range(n)])
        want = oracle_rep(oracle, c, X, beta[s:s + 1], th[:, s:s + 1], nn, m,
                          0, z)
        assert rel(got[:, s], want[:, 0]) < RTOL


def test_replicates_reject_a_neighbor_that_is_not_a_predecessor(oracle,
                                                                 nngp):
    n, m, p, S = 200, 4, 1, 2
    c, X, beta, th = problem(n, m, p, S, 0, 3)
    nn = oracle.mkNNIndxCB(c, m)
    bad = nn["nnIndx"].copy()
    lu = nn["nnIndxLU"]
    bad[lu[50]] = 120  # site 50 pointing at a later site
    with pytest.raises(nngp.NngpError):
        nngp.rNNGPReplicated(X, p, n, m, c, 0, bad, lu, beta, th, S)
