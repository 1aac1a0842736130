"""Pins the CPU oracle (oracle/) against facts that do not come from it:
closed forms, scipy, brute force, and dense-Gaussian identities.

Every check is built so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.special as sp
from scipy.spatial import cKDTree

from conftest import uniform_coords


# ---------------------------------------------------------------------------
# layout (src/util.cpp:55-70)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,m", [(2, 1), (10, 3), (57, 10), (300, 15),
                                 (40, 39)])
def test_layout_is_contiguous_and_complete(oracle, n, m):
    offs, cnts = zip(*[oracle.getNNIndx(i, m) for i in range(n)])
    assert cnts[0] == 0
    for i in range(1, n):
        assert cnts[i] == min(i, m)
    # site blocks tile [0, nIndx) without gaps or overlaps
    covered = np.zeros(oracle.n_indx(n, m), dtype=int)
    for i in range(1, n):
        covered[offs[i]:offs[i] + cnts[i]] += 1
    assert (covered == 1).all()


# ---------------------------------------------------------------------------
# special functions and correlations (src/util.cpp:208-245)
# ---------------------------------------------------------------------------
def test_bessel_k_matches_scipy(oracle):
    worst = 0.0
    for nu in [0.0, 0.01, 0.1, 0.3, 0.5, 0.77, 1.0, 1.5, 2.0, 2.5, 3.3, 5.0]:
        for x in np.concatenate([np.logspace(-4, np.log10(60), 50),
                                 [1.999, 2.0, 2.001]]):
            a, b = oracle.bessel_k(x, nu), sp.kv(nu, x)
            worst = max(worst, abs(a - b) / abs(b))
    assert worst < 1e-13


def test_bessel_k_half_integer_closed_forms(oracle):
    for x in [1e-3, 0.2, 1.0, 2.0, 3.5, 20.0]:
        k05 = np.sqrt(np.pi / (2 * x)) * np.exp(-x)
        assert oracle.bessel_k(x, 0.5) == pytest.approx(k05, rel=1e-14)
        assert oracle.bessel_k(x, 1.5) == pytest.approx(
            k05 * (1 + 1 / x), rel=1e-14)
        assert oracle.bessel_k(x, 2.5) == pytest.approx(
            k05 * (1 + 3 / x + 3 / x ** 2), rel=1e-14)


@pytest.mark.parametrize("phi", [0.5, 6.0, 40.0])
def test_correlation_models(oracle, phi):
    for D in [0.0, 1e-4, 0.01, 0.1, 0.5, 1.0 / phi, 0.9 / phi, 1.3 / phi]:
        t = phi * D
        assert oracle.spCor(D, phi, 0, 0) == pytest.approx(np.exp(-t),
                                                           rel=1e-15)
        assert oracle.spCor(D, phi, 0, 3) == pytest.approx(np.exp(-t * t),
                                                           rel=1e-15)
        sph = 1.0 if D == 0 else (1 - 1.5 * t + 0.5 * t ** 3 if D <= 1 / phi
                                  else 0.0)
        assert oracle.spCor(D, phi, 0, 1) == pytest.approx(sph, abs=1e-15)
        # matern at half-integer nu reduces to closed forms
        assert oracle.spCor(D, phi, 0.5, 2) == pytest.approx(np.exp(-t),
                                                             rel=1e-13)
        assert oracle.spCor(D, phi, 1.5, 2) == pytest.approx(
            (1 + t) * np.exp(-t), rel=1e-13)
        assert oracle.spCor(D, phi, 2.5, 2) == pytest.approx(
            (1 + t + t * t / 3) * np.exp(-t), rel=1e-13)
    # spherical support ends exactly at D = 1/phi
    assert oracle.spCor(1.0 / phi, phi, 0, 1) == pytest.approx(0.0, abs=1e-15)
    assert oracle.spCor(2.0 / phi, phi, 0, 1) == 0.0


def test_matern_general_nu_matches_scipy(oracle):
    for nu in [0.3, 0.8, 1.2, 2.1, 3.7]:
        for x in [1e-3, 0.05, 0.7, 1.9, 2.1, 6.0, 25.0]:
            ref = x ** nu / (2 ** (nu - 1) * sp.gamma(nu)) * sp.kv(nu, x)
            assert oracle.spCor(x / 4.0, 4.0, nu, 2) == pytest.approx(
                ref, rel=2e-13)


def test_bad_cov_model_is_nan(oracle):
    assert np.isnan(oracle.spCor(0.1, 1.0, 0.5, 7))


# ---------------------------------------------------------------------------
# sorting and neighbor search (src/nn.cpp)
# ---------------------------------------------------------------------------
def test_rsort_with_index_sorts_pairs(oracle):
    rng = np.random.default_rng(0)
    for n in [1, 2, 9, 10, 15, 40, 200]:
        x = rng.normal(size=n)
        idx = np.arange(n, dtype=np.int32)
        xs, ix = oracle.rsort_with_index(x, idx)
        assert np.array_equal(xs, np.sort(x))
        assert np.array_equal(x[ix], xs)
    xs, ix = oracle.rsort_with_index(np.array([3.0, np.nan, 1.0, 2.0]),
                                     np.arange(4, dtype=np.int32))
    assert np.array_equal(xs[:3], [1.0, 2.0, 3.0]) and np.isnan(xs[3])
    assert list(ix) == [2, 3, 0, 1]


def test_rsort_with_index_shell_order_of_ties_known_answer(oracle):
    # Traced by hand through R's shell sort (gaps 4 then 1 for n = 10): the
    # h = 4 pass moves the 3.0 at position 5 past the 1.5 to position 9,
    # and the stable h = 1 pass keeps it behind the other 3.0s, so the tie
    # order is not the input order.
    x = np.array([0.0, 1.0, 2.0, 3.0, 3.0, 3.0, 3.0, 3.0, 3.0, 1.5])
    xs, ix = oracle.rsort_with_index(x, np.arange(10, dtype=np.int32))
    assert list(ix) == [0, 1, 9, 2, 3, 4, 6, 7, 8, 5]
    # below n = 9 there is only the h = 1 pass: a stable insertion sort
    x = np.array([2.0, 1.0, 2.0, 1.0, 0.5, 2.0, 1.0, 0.5])
    xs, ix = oracle.rsort_with_index(x, np.arange(8, dtype=np.int32))
    assert list(ix) == list(np.argsort(x, kind="stable"))


def test_knn_ties_on_a_2x2_lattice_known_answer(oracle):
    # Sites (x order): s0 (0,0), s1 (0,1), s2 (1,0), s3 (1,1); m = 2.
    # Traced by hand through src/nn.cpp.  Code book: u = (0, 1, 1, 2) keeps
    # s1 before s2 (gap-1 sort); s3's walk goes down only (it is last in u),
    # meets s2 first, then s1 at the same distance 1 (inserted: 1 < inf),
    # then stops at s0 ((du)^2 = 4 > 2).  The stable sort leaves s2 first.
    # The scan meets s1 before s2, so it lists s1 first.
    c = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0], [1.0, 1.0]])
    cb = oracle.mkNNIndxCB(c, 2)
    br = oracle.mkNNIndx(c, 2)
    assert list(cb["nnIndxLU"]) == [0, 0, 1, 3, 0, 1, 2, 2]
    assert list(cb["nnIndx"]) == [0, 0, 1, 2, 1]
    assert list(br["nnIndx"]) == [0, 0, 1, 1, 2]
    r2 = np.sqrt(2.0)
    assert list(cb["nnDist"]) == [1.0, 1.0, r2, 1.0, 1.0]
    assert list(br["nnDist"]) == [1.0, 1.0, r2, 1.0, 1.0]


@pytest.mark.parametrize("kind", ["lattice", "rounded", "duplicates"])
def test_knn_tied_inputs_keep_the_m_smallest_distances(oracle, kind):
    # Whatever the tie order, each list holds the m smallest distances
    # among the predecessors (the walk's pruning is exact up to ties).
    rng = np.random.default_rng(7)
    if kind == "lattice":
        g = np.arange(25) / 25.0
        c = np.array([(a, b) for a in g for b in g])
    elif kind == "rounded":
        c = np.round(rng.uniform(size=(700, 2)), 2)
    else:
        c = rng.uniform(size=(500, 2))
        c = np.vstack([c, c[:80], c[:10]])
    c = c[np.argsort(c[:, 0], kind="stable")]
    n, m = c.shape[0], 10
    for fn in (oracle.mkNNIndx, oracle.mkNNIndxCB):
        r = fn(c, m)
        lu = r["nnIndxLU"]
        for i in range(1, n):
            off, cnt = lu[i], lu[n + i]
            dx = c[i, 0] - c[:i, 0]
            dy = c[i, 1] - c[:i, 1]
            best = np.sort(np.sqrt(dx * dx + dy * dy))[:cnt]
            got = r["nnDist"][off:off + cnt]
            assert np.array_equal(got, best), (fn.__name__, i)
            j = r["nnIndx"][off:off + cnt]
            assert len(set(j)) == cnt and np.all(j < i)


def brute_pred_knn(coords, m):
    """m nearest predecessors by the reference's squared-distance arithmetic,
    written directly in numpy."""
    n = coords.shape[0]
    out = []
    for i in range(1, n):
        dx = coords[i, 0] - coords[:i, 0]
        dy = coords[i, 1] - coords[:i, 1]
        d2 = dx * dx + dy * dy
        k = min(i, m)
        order = np.argsort(d2, kind="stable")[:k]
        out.append((order, np.sqrt(d2[order])))
    return out


@pytest.mark.parametrize("n,m,seed", [(300, 5, 1), (600, 10, 2),
                                      (400, 15, 3)])
def test_knn_brute_and_cb_match_numpy(oracle, n, m, seed):
    c = uniform_coords(n, seed)
    ref = brute_pred_knn(c, m)
    for fn in (oracle.mkNNIndx, oracle.mkNNIndxCB):
        r = fn(c, m)
        lu = r["nnIndxLU"]
        for i in range(1, n):
            off, cnt = lu[i], lu[n + i]
            assert np.array_equal(r["nnIndx"][off:off + cnt], ref[i - 1][0])
            np.testing.assert_allclose(r["nnDist"][off:off + cnt],
                                       ref[i - 1][1], rtol=0, atol=0)


def test_knn_cb_matches_kdtree_on_prefixes(oracle):
    n, m = 2000, 10
    c = uniform_coords(n, 11)
    r = oracle.mkNNIndxCB(c, m)
    lu = r["nnIndxLU"]
    for i in [11, 50, 333, 999, 1500, 1999]:
        d, j = cKDTree(c[:i]).query(c[i], k=m)
        off = lu[i]
        assert set(r["nnIndx"][off:off + m]) == set(j)
        np.testing.assert_allclose(r["nnDist"][off:off + m], d, rtol=1e-15)


def test_knn_custom_order(oracle):
    # NNGP order different from x order: predecessors are by index, not x
    rng = np.random.default_rng(5)
    c = rng.uniform(size=(400, 2))
    ref = brute_pred_knn(c, 8)
    r = oracle.mkNNIndxCB(c, 8)
    lu = r["nnIndxLU"]
    for i in range(1, 400):
        off, cnt = lu[i], lu[400 + i]
        assert np.array_equal(r["nnIndx"][off:off + cnt], ref[i - 1][0])


def test_nn2_orders_ties_by_ascending_row(oracle):
    g = np.arange(20) / 20.0
    c = np.array([(a, b) for a in g for b in g])[::-1].copy()
    c0 = np.vstack([c[::7] + 0.025, c[::11]])  # cell centres and sites
    k = 9
    idx, dist = oracle.nn2(c, c0, k)
    for i in range(c0.shape[0]):
        dx = c0[i, 0] - c[:, 0]
        dy = c0[i, 1] - c[:, 1]
        d2 = dx * dx + dy * dy
        order = np.lexsort((np.arange(c.shape[0]), d2))[:k]
        assert np.array_equal(idx[i], order)
        assert np.array_equal(dist[i], np.sqrt(d2[order]))


def test_nn2_matches_kdtree(oracle):
    rng = np.random.default_rng(9)
    c = rng.uniform(size=(500, 2))
    c0 = rng.uniform(size=(60, 2))
    idx, dist = oracle.nn2(c, c0, 7)
    d, j = cKDTree(c).query(c0, k=7)
    assert np.array_equal(idx, j)
    np.testing.assert_allclose(dist, d, rtol=1e-15)


# ---------------------------------------------------------------------------
# dense kernels
# ---------------------------------------------------------------------------
def test_potrf_potri_symv_match_lapack(oracle):
    rng = np.random.default_rng(3)
    for n in [1, 2, 5, 15, 32]:
        A = rng.normal(size=(n, n))
        A = A @ A.T + n * np.eye(n)
        L, info = oracle.potrf_L(A)
        assert info == 0
        np.testing.assert_allclose(np.tril(L), np.linalg.cholesky(A),
                                   rtol=1e-13, atol=1e-13)
        Ai, info = oracle.potri_L(np.tril(L))
        assert info == 0
        np.testing.assert_allclose(np.tril(Ai), np.tril(np.linalg.inv(A)),
                                   rtol=1e-11, atol=1e-13)
        x = rng.normal(size=n)
        full = np.tril(Ai) + np.tril(Ai, -1).T
        np.testing.assert_allclose(oracle.symv_L(Ai, x), full @ x,
                                   rtol=1e-12, atol=1e-13)
    # not positive definite: 1-based failing column
    A = np.array([[1.0, 2.0], [2.0, 1.0]])
    assert oracle.potrf_L(A)[1] == 2


# ---------------------------------------------------------------------------
# updateBF / Q: exact-GP identity and residual checks
# ---------------------------------------------------------------------------
def cov_dense(c, s2, t2, phi, nu, cov, oracle):
    n = c.shape[0]
    D = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    R = np.vectorize(lambda d: oracle.spCor(d, phi, nu, cov))(D)
    return s2 * R + t2 * np.eye(n)


def nngp_precision(B, F, nnIndx, lu, n):
    """Dense (I - B)^T F^-1 (I - B) from the flat layout."""
    IB = np.eye(n)
    for i in range(n):
        off, cnt = lu[i], lu[n + i]
        for j in range(cnt):
            IB[i, nnIndx[off + j]] -= B[off + j]
    return IB.T @ np.diag(1.0 / F) @ IB


@pytest.mark.parametrize("cov,nu", [(0, 0.0), (1, 0.0), (2, 1.3), (3, 0.0)])
def test_updateBF_exact_gp_identity(oracle, cov, nu):
    """With every predecessor as neighbor (m = n-1) the NNGP is the full GP:
    sum log F = log det C and the NNGP precision equals C^-1."""
    n = 40
    c = uniform_coords(n, 21)
    phi = 3.0 if cov != 1 else 0.8
    s2, t2 = 2.0, 0.3
    nn = oracle.mkNNIndx(c, n - 1)
    flat = c.T.reshape(-1).copy()
    B, F, ld = oracle.updateBF(flat, nn["nnIndx"], nn["nnIndxLU"], n, n - 1,
                               [s2, t2, phi, nu], cov)
    C = cov_dense(c, s2, t2, phi, nu, cov, oracle)
    sign, logdet = np.linalg.slogdet(C)
    assert sign > 0
    assert ld == pytest.approx(logdet, rel=1e-11)
    P = nngp_precision(B, F, nn["nnIndx"], nn["nnIndxLU"], n)
    np.testing.assert_allclose(P, np.linalg.inv(C), rtol=1e-8, atol=1e-8)


def test_updateBF_solves_each_system(oracle):
    n, m = 500, 10
    c = uniform_coords(n, 7)
    nn = oracle.mkNNIndxCB(c, m)
    s2, t2, phi = 5.0, 1.0, 6.0
    B, F, ld = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                               nn["nnIndxLU"], n, m, [s2, t2, phi], 0)
    lu = nn["nnIndxLU"]
    for i in [1, 2, 9, 10, 11, 250, 499]:
        off, cnt = lu[i], lu[n + i]
        N = nn["nnIndx"][off:off + cnt]
        pts = c[N]
        Dn = np.sqrt(((pts[:, None] - pts[None]) ** 2).sum(-1))
        Cn = s2 * np.exp(-phi * Dn) + t2 * np.eye(cnt)
        cv = s2 * np.exp(-phi * np.sqrt(((pts - c[i]) ** 2).sum(-1)))
        b = B[off:off + cnt]
        assert np.linalg.norm(Cn @ b - cv) <= 1e-12 * np.linalg.norm(cv)
        assert F[i] == pytest.approx(s2 + t2 - b @ cv, rel=1e-12)
    assert F[0] == s2 + t2
    assert ld == pytest.approx(np.log(F).sum(), rel=1e-14)


def test_Q_equals_dense_quadratic_form(oracle):
    n, m = 120, 6
    c = uniform_coords(n, 8)
    nn = oracle.mkNNIndxCB(c, m)
    B, F, _ = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                              nn["nnIndxLU"], n, m, [1.0, 0.2, 4.0], 0)
    rng = np.random.default_rng(1)
    u, v = rng.normal(size=n), rng.normal(size=n)
    P = nngp_precision(B, F, nn["nnIndx"], nn["nnIndxLU"], n)
    assert oracle.Q(B, F, u, v, nn["nnIndx"], nn["nnIndxLU"]) == \
        pytest.approx(u @ P @ v, rel=1e-11)
    # the quadratic form is symmetric in (u, v)
    assert oracle.Q(B, F, u, v, nn["nnIndx"], nn["nnIndxLU"]) == \
        pytest.approx(oracle.Q(B, F, v, u, nn["nnIndx"], nn["nnIndxLU"]),
                      rel=1e-13)


def test_nngp_draw_has_nngp_covariance(oracle):
    # w = (I - B)^-1 F^1/2 z, so the draw map is linear with that matrix
    n, m = 60, 5
    c = uniform_coords(n, 4)
    nn = oracle.mkNNIndxCB(c, m)
    B, F, _ = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                              nn["nnIndxLU"], n, m, [1.0, 0.0, 5.0], 0)
    IB = np.eye(n)
    lu = nn["nnIndxLU"]
    for i in range(n):
        for j in range(lu[n + i]):
            IB[i, nn["nnIndx"][lu[i] + j]] -= B[lu[i] + j]
    z = np.random.default_rng(2).normal(size=n)
    w = oracle.nngp_draw(B, F, nn["nnIndx"], nn["nnIndxLU"], z)
    np.testing.assert_allclose(w, np.linalg.solve(IB, np.sqrt(F) * z),
                               rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------------------
# the kriging server's counter-based normals
# ---------------------------------------------------------------------------
def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    M = 0xFFFFFFFF
    cases = [
        ((0, 0, 0, 0), (0, 0),
         (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((M, M, M, M), (M, M),
         (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
         (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in cases:
        assert tuple(oracle.philox4x32_10(ctr, key)) == want


def test_philox_normals_are_standard_normal(oracle):
    import scipy.stats as ss
    z = oracle.philox_normals(0xDEADBEEF12345678, 200, 100)
    assert ss.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 4 / np.sqrt(z.size)
    assert abs(z.var() - 1) < 6 * np.sqrt(2 / z.size)
    # neighboring counters and seeds are uncorrelated
    zz = z.reshape(200, 100)
    assert abs(np.corrcoef(zz[:, 0], zz[:, 1])[0, 1]) < 0.3
    z2 = oracle.philox_normals(0xDEADBEEF12345679, 200, 100)
    assert abs(np.corrcoef(z, z2)[0, 1]) < 0.05
