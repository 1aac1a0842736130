"""GPU parity: libnngp_hip.so (through its C ABI) against the CPU oracle on
the same seeded inputs.

Bars: neighbor indices and distances bit-exact; B, F, log-det, Q, kriging
and conjugate outputs within 1e-10 relative (FP64 path); MCMC chains with
the same host random numbers accept and reject identically, samples within
1e-9 relative.
"""
import time

import numpy as np
import pytest

from conftest import uniform_coords

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def rel_norm_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def per_site_B_err(B1, B0, lu, n):
    worst = 0.0
    for i in range(1, n):
        off, cnt = lu[i], lu[n + i]
        b0 = B0[off:off + cnt]
        d = np.max(np.abs(B1[off:off + cnt] - b0)) / max(np.max(np.abs(b0)),
                                                         1e-300)
        worst = max(worst, d)
    return worst


# ---------------------------------------------------------------------------
# neighbor search
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,m,seed", [(2000, 10, 1), (5000, 15, 2),
                                      (3000, 32, 3), (1500, 1, 4),
                                      (700, 64, 5)])
def test_knn_bit_exact(oracle, nngp, n, m, seed):
    print(f"knn n={n} m={m} seed={seed}", flush=True)
    c = uniform_coords(n, seed)
    ref = oracle.mkNNIndx(c, m)
    got = nngp.mkNNIndxCB(c, m)
    assert np.array_equal(got["nnIndxLU"], ref["nnIndxLU"])
    assert np.array_equal(got["nnIndx"], ref["nnIndx"])
    assert np.array_equal(got["nnDist"], ref["nnDist"])  # bit for bit


def test_knn_custom_order_bit_exact(oracle, nngp):
    rng = np.random.default_rng(8)
    c = rng.uniform(size=(2500, 2))  # not x-ordered
    ref = oracle.mkNNIndx(c, 12)
    got = nngp.mkNNIndx(c, 12)
    assert np.array_equal(got["nnIndx"], ref["nnIndx"])
    assert np.array_equal(got["nnDist"], ref["nnDist"])


def test_knn_clustered_anisotropic(oracle, nngp):
    rng = np.random.default_rng(12)
    cen = rng.uniform(size=(10, 2))
    pts = cen[rng.integers(0, 10, 3000)] + rng.normal(scale=0.02,
                                                      size=(3000, 2))
    pts[:, 0] *= 3.0
    pts = pts[np.argsort(pts[:, 0], kind="stable")]
    ref = oracle.mkNNIndx(pts, 20)
    got = nngp.mkNNIndxCB(pts, 20)
    assert np.array_equal(got["nnIndx"], ref["nnIndx"])
    assert np.array_equal(got["nnDist"], ref["nnDist"])


@pytest.mark.parametrize("k", [1, 10, 15, 40])
def test_nn2_bit_exact(oracle, nngp, k):
    rng = np.random.default_rng(20 + k)
    c = rng.uniform(size=(4000, 2))
    c0 = rng.uniform(size=(900, 2))
    i_ref, d_ref = oracle.nn2(c, c0, k)
    i_got, d_got = nngp.nn2(c, c0, k)
    assert np.array_equal(i_got, i_ref)
    assert np.array_equal(d_got, d_ref)


# ---------------------------------------------------------------------------
# updateBF / updateConjBF / Q
# ---------------------------------------------------------------------------
CASES = [
    # (n, m, cov, theta) -- per-lane kernel for cov != 2 and m <= 16,
    # wavefront kernel for matern or m > 16
    (2000, 10, 0, [5.0, 1.0, 6.0, 0.0]),
    (3000, 15, 0, [2.0, 0.2, 12.0, 0.0]),
    (1500, 12, 1, [1.0, 0.1, 3.0, 0.0]),
    (1500, 8, 3, [1.5, 0.5, 4.0, 0.0]),
    (1200, 15, 2, [1.0, 0.1, 8.0, 1.5]),
    (800, 10, 2, [1.0, 0.05, 20.0, 2.5]),
    (900, 32, 0, [1.0, 0.05, 10.0, 0.0]),
    (600, 32, 2, [1.0, 0.05, 20.0, 0.7]),
    (500, 48, 0, [1.0, 0.1, 5.0, 0.0]),
    (300, 64, 2, [1.0, 0.2, 4.0, 1.2]),
    (400, 3, 0, [1.0, 0.5, 2.0, 0.0]),
]


@pytest.mark.parametrize("n,m,cov,theta", CASES)
def test_updateBF_matches_oracle(oracle, nngp, n, m, cov, theta):
    print(f"updateBF n={n} m={m} cov={cov} theta={theta}", flush=True)
    c = uniform_coords(n, 100 + n + m)
    nn = oracle.mkNNIndx(c, m)
    flat = c.T.reshape(-1).copy()
    B0, F0, ld0 = oracle.updateBF(flat, nn["nnIndx"], nn["nnIndxLU"], n, m,
                                  theta, cov)
    B1, F1, ld1 = nngp.updateBF(c, nn["nnIndx"], nn["nnIndxLU"], m, theta,
                                cov)
    assert per_site_B_err(B1, B0, nn["nnIndxLU"], n) < RTOL
    assert rel_err(F1, F0) < RTOL
    assert abs(ld1 - ld0) / abs(ld0) < RTOL


SHAPES = [(0, 64, 0), (0, 128, 0), (0, 256, 0), (0, 64, 2), (0, 128, 2),
          (0, 256, 2), (1, 64, 0), (1, 128, 0), (1, 256, 0), (2, 64, 0),
          (2, 128, 0), (2, 256, 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("cov,m,theta", [(0, 15, [2.0, 0.2, 12.0]),
                                         (2, 10, [1.0, 0.3, 6.0, 1.5]),
                                         (1, 6, [1.5, 0.5, 2.0])])
def test_every_launch_shape_matches_oracle(oracle, nngp, cov, m, theta):
    """Each updateBF launch shape the autotuner may pick (one site per lane
    in 64/128/256-thread blocks, one per wavefront in 1/2/4-wave blocks, one
    per 16-lane row in 64/128/256-thread blocks) gives B, F, the
    log-determinant and Q within the parity bound; n is not a multiple of
    64, so the last block is ragged."""
    n = 5000 + 37
    print(f"shapes cov={cov} m={m}", flush=True)
    c = uniform_coords(n, 300 + m)
    rng = np.random.default_rng(m)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, -1.0] + rng.normal(size=n)
    nn = oracle.mkNNIndx(c, m)
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, theta, cov)
    beta = np.array([1.0, -1.0])
    r = X @ beta - y
    q0 = oracle.Q(B0, F0, r, r, nn["nnIndx"], nn["nnIndxLU"])
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, cov, X, y) as md:
        for fam, th, mode in SHAPES:
            md.set_shape(fam, th, mode)
            out = md.update(theta, slot=0, beta=beta)
            B, F = md.get_bf(0)
            assert per_site_B_err(B, B0, nn["nnIndxLU"], n) < RTOL
            assert rel_err(F, F0) < RTOL
            assert abs(out["logDet"] - ld0) < RTOL * abs(ld0)
            assert abs(out["qResid"] - q0) < RTOL * abs(q0)


@pytest.mark.gpu
@pytest.mark.parametrize("cov,m,p,theta", [
    (2, 32, 3, [1.0, 0.05, 20.0, 0.7]),
    (2, 24, 1, [2.0, 0.2, 8.0, 1.2]),
    (2, 17, 0, [1.0, 0.3, 6.0, 2.2]),
    (2, 16, 2, [1.0, 0.1, 9.0, 0.35]),
    (0, 32, 2, [1.0, 0.05, 10.0]),
    (3, 20, 3, [1.0, 0.1, 5.0]),
    (1, 28, 1, [1.5, 0.5, 2.0]),
])
def test_row_kernel_matches_oracle_and_the_wavefront_kernel(oracle, nngp,
                                                            cov, m, p, theta):
    """The 16-lane-row kernel (four sites a wavefront; lane u of a site's
    row holds rows u and u + 16 of its system, DPP broadcasts within the
    row) at 64, 128 and 256 threads: B, F, the log determinant and Q within
    the parity bound of the oracle, and the cross-products within it of the
    wavefront kernel's.  n = 4003 leaves a ragged last group of four, and
    the first m sites have fewer than m neighbors (identity rows)."""
    n = 4003
    print(f"row kernel cov={cov} m={m} p={p}", flush=True)
    c = uniform_coords(n, 700 + m + p)
    rng = np.random.default_rng(m + 100 * p)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))]) \
        if p else None
    beta = np.linspace(1.0, -1.0, p) if p else np.zeros(1)
    y = (X @ beta if p else 0.0) + rng.normal(size=n)
    nn = oracle.mkNNIndx(c, m)
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, theta, cov)
    if p:
        r = X @ beta - y
        q0 = oracle.Q(B0, F0, r, r, nn["nnIndx"], nn["nnIndxLU"])
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, cov, X, y) as md:
        md.set_shape(1, 256, 0)
        ref = md.update(theta, slot=0, beta=beta)
        for th in (64, 128, 256):
            md.set_shape(2, th, 0)
            out = md.update(theta, slot=1, beta=beta)
            B, F = md.get_bf(1)
            assert per_site_B_err(B, B0, nn["nnIndxLU"], n) < RTOL
            assert rel_err(F, F0) < RTOL
            assert abs(out["logDet"] - ld0) < RTOL * abs(ld0)
            if p:
                assert abs(out["qResid"] - q0) < RTOL * abs(q0)
                for k in ("XtViX", "XtViy"):
                    assert np.max(np.abs(out[k] - ref[k])) \
                        < RTOL * np.max(np.abs(ref[k]))


@pytest.mark.gpu
def test_row_kernel_shapes_are_checked(nngp):
    """The row kernel takes m <= 32 and FP64 passes only."""
    c = uniform_coords(600, 10)
    for m, ok in ((32, True), (33, False)):
        nn = nngp.mkNNIndxCB(c, m)
        with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, 0) as md:
            if ok:
                md.set_shape(2, 128, 0)
            else:
                with pytest.raises(nngp.NngpError):
                    md.set_shape(2, 128, 0)
    nn = nngp.mkNNIndxCB(c, 15)
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], 15, 2,
                    precision=1) as md:
        with pytest.raises(nngp.NngpError):
            md.set_shape(2, 64, 0)


def test_updateBF_with_a_non_canonical_layout(oracle, nngp):
    """nnIndxLU need not follow the closed form of src/util.cpp:55-70: with
    the sites' neighbor blocks stored in reverse order the kernels read the
    offsets and counts instead of computing them, with the same result."""
    n, m, theta = 3000, 12, [2.0, 0.3, 9.0]
    c = uniform_coords(n, 41)
    nn = oracle.mkNNIndx(c, m)
    idx, lu = nn["nnIndx"], nn["nnIndxLU"]
    tot = len(idx)
    idx2 = np.zeros_like(idx)
    lu2 = lu.copy()
    for i in range(n):
        o, k = lu[i], lu[n + i]
        lu2[i] = tot - (o + k)
        idx2[lu2[i]:lu2[i] + k] = idx[o:o + k]
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), idx2, lu2, n, m,
                                  theta, 0)
    B1, F1, ld1 = nngp.updateBF(c, idx2, lu2, m, theta, 0)
    assert per_site_B_err(B1, B0, lu2, n) < RTOL
    assert rel_err(F1, F0) < RTOL
    assert abs(ld1 - ld0) < RTOL * abs(ld0)


@pytest.mark.gpu
@pytest.mark.parametrize("cov,m,p,theta", [
    (0, 15, 3, [2.0, 0.2, 12.0]),
    (0, 16, 2, [2.0, 0.2, 12.0]),
    (1, 12, 1, [1.5, 0.5, 2.0]),
    (3, 10, 0, [1.0, 0.3, 5.0]),
    (2, 15, 1, [1.0, 0.3, 6.0, 1.5]),
    (2, 12, 3, [1.0, 0.3, 6.0, 2.5]),
])
def test_lane_lds_staging_is_bitwise_the_plain_lane_kernel(oracle, nngp, cov,
                                                           m, p, theta):
    """Mode 2 of the per-lane kernel stages the covariate rows in LDS with
    direct loads and adds the same terms in the same order: B, F, the log
    determinant and every cross-product equal mode 0's bit for bit, for row
    strides of 2 (p <= 1) and 4 doubles (p = 2, 3), at 64 and 128 threads,
    on a ragged last block; both match the oracle."""
    n = 4000 + 21
    print(f"lane lds cov={cov} m={m} p={p}", flush=True)
    c = uniform_coords(n, 500 + m + p)
    rng = np.random.default_rng(m + 10 * p)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))]) \
        if p else None
    beta = np.linspace(1.0, -1.0, p) if p else np.zeros(1)
    y = (X @ beta if p else 0.0) + rng.normal(size=n)
    nn = oracle.mkNNIndxCB(c, m)
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, theta, cov)
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, cov, X, y) as md:
        for th in (64, 128):
            res = []
            for mode in (0, 2):
                md.set_shape(0, th, mode)
                out = md.update(theta, slot=mode // 2, beta=beta)
                res.append((out, *md.get_bf(mode // 2)))
            (o0, Ba, Fa), (o2, Bb, Fb) = res
            assert np.array_equal(Ba, Bb) and np.array_equal(Fa, Fb)
            assert o0["logDet"] == o2["logDet"]
            assert o0["qResid"] == o2["qResid"]
            if p:
                assert np.array_equal(o0["XtViX"], o2["XtViX"])
                assert np.array_equal(o0["XtViy"], o2["XtViy"])
                assert o0["ytViy"] == o2["ytViy"]
    assert per_site_B_err(Bb, B0, nn["nnIndxLU"], n) < RTOL
    assert rel_err(Fb, F0) < RTOL
    assert abs(o2["logDet"] - ld0) < RTOL * abs(ld0)


@pytest.mark.gpu
@pytest.mark.parametrize("cov,theta", [(0, [2.0, 0.2, 12.0e-4]),
                                       (3, [2.0, 0.2, 3.0e-4])])
def test_projected_coordinates_keep_parity(oracle, nngp, cov, theta):
    """Sites in a projected system: a 1e4 x 1e4 window at an offset of
    (5e5, 4e6), where the coordinates carry 10 more significant digits than
    the distances.  The per-lane kernels form neighbor offsets from the site
    before scaling (the exponential and gaussian run in base-2 units), so B,
    F and the log determinant keep the parity bound in both launch modes."""
    n, m = 3000 + 5, 15
    print(f"projected cov={cov}", flush=True)
    c = uniform_coords(n, 808) * 1e4 + np.array([5e5, 4e6])
    c = c[np.argsort(c[:, 0], kind="stable")]
    nn = oracle.mkNNIndxCB(c, m)
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, theta, cov)
    rng = np.random.default_rng(8)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ [1.0, 0.5, -1.0] + rng.normal(size=n)
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, cov, X, y) as md:
        for mode in (0, 2):
            md.set_shape(0, 64, mode)
            out = md.update(theta, slot=0, beta=[1.0, 0.5, -1.0])
            B, F = md.get_bf(0)
            assert per_site_B_err(B, B0, nn["nnIndxLU"], n) < RTOL
            assert rel_err(F, F0) < RTOL
            assert abs(out["logDet"] - ld0) < RTOL * abs(ld0)


@pytest.mark.gpu
def test_set_shape_rejects_shapes_that_do_not_apply(nngp):
    n, m = 500, 20
    c = uniform_coords(n, 9)
    nn = nngp.mkNNIndxCB(c, m)
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, 0) as md:
        with pytest.raises(nngp.NngpError):
            md.set_shape(0, 64, 0)  # m > 16: no per-lane kernel
        with pytest.raises(nngp.NngpError):
            md.set_shape(1, 96, 0)
        md.set_shape(1, 128, 0)
    # modes of the other precision: 1 and 3 are the mixed kernel's
    # covariance copies, 2 the FP64 kernel's staged rows
    nn15 = nngp.mkNNIndxCB(c, 15)
    with nngp.Model(c, nn15["nnIndx"], nn15["nnIndxLU"], 15, 0) as md:
        for mode in (1, 3):
            with pytest.raises(nngp.NngpError):
                md.set_shape(0, 64, mode)
        md.set_shape(0, 64, 2)
    with nngp.Model(c, nn15["nnIndx"], nn15["nnIndxLU"], 15, 0,
                    precision=1) as md:
        with pytest.raises(nngp.NngpError):
            md.set_shape(0, 64, 2)
        with pytest.raises(nngp.NngpError):
            md.set_shape(2, 256, 0)  # the row kernel is FP64 only
        for mode in (0, 1, 3):
            md.set_shape(0, 64, mode)


@pytest.mark.parametrize("cov,nu", [(0, 0.0), (2, 1.5), (3, 0.0)])
def test_updateConjBF_matches_oracle(oracle, nngp, cov, nu):
    n, m = 1500, 15
    c = uniform_coords(n, 77)
    nn = oracle.mkNNIndx(c, m)
    B0, F0 = oracle.updateConjBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                 nn["nnIndxLU"], n, m, 8.0, 0.1, nu, cov)
    B1, F1 = nngp.updateConjBF(c, nn["nnIndx"], nn["nnIndxLU"], m, 8.0, 0.1,
                               nu, cov)
    assert per_site_B_err(B1, B0, nn["nnIndxLU"], n) < RTOL
    assert rel_err(F1, F0) < RTOL


def test_Q_matches_oracle(oracle, nngp):
    n, m = 4000, 15
    c = uniform_coords(n, 5)
    nn = oracle.mkNNIndx(c, m)
    B, F, _ = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                              nn["nnIndxLU"], n, m, [2.0, 0.2, 12.0], 0)
    rng = np.random.default_rng(0)
    u, v = rng.normal(size=n), rng.normal(size=n)
    q0 = oracle.Q(B, F, u, v, nn["nnIndx"], nn["nnIndxLU"])
    q1 = nngp.Q(B, F, u, v, nn["nnIndx"], nn["nnIndxLU"], m)
    assert abs(q1 - q0) / abs(q0) < RTOL


def test_model_pass_cross_products_match_oracle_Q(oracle, nngp):
    """The fused pass's X^T V^-1 X, X^T V^-1 y, y^T V^-1 y and Q(r, r)
    against the oracle's Q on the same B, F (src/rNNGP.cpp:281-286)."""
    n, m, p = 3000, 10, 3
    c = uniform_coords(n, 31)
    rng = np.random.default_rng(31)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))])
    y = X @ [1.0, 2.0, -1.0] + rng.normal(size=n)
    nn = oracle.mkNNIndx(c, m)
    theta = [2.0, 0.3, 9.0]
    B, F, ld = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                               nn["nnIndxLU"], n, m, theta, 0)
    beta = np.array([0.9, 2.1, -0.8])
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, 0, X, y) as md:
        out = md.update(theta, slot=1, beta=beta)
        q_again = md.qresid(beta, slot=1)
        B1, F1 = md.get_bf(1)
    L, lu = nn["nnIndx"], nn["nnIndxLU"]
    Q = lambda u, v: oracle.Q(B, F, u, v, L, lu)
    r = X @ beta - y
    assert abs(out["logDet"] - ld) / abs(ld) < RTOL
    assert abs(out["qResid"] - Q(r, r)) / Q(r, r) < RTOL
    assert q_again == out["qResid"]
    for a in range(p):
        assert abs(out["XtViy"][a] - Q(X[:, a], y)) <= RTOL * abs(Q(y, y))
        for b in range(p):
            ref = Q(X[:, a], X[:, b])
            assert abs(out["XtViX"][a, b] - ref) <= RTOL * Q(X[:, a], X[:, a])
    assert abs(out["ytViy"] - Q(y, y)) / Q(y, y) < RTOL
    assert per_site_B_err(B1, B, lu, n) < RTOL


# ---------------------------------------------------------------------------
# rNNGP chain identity
# ---------------------------------------------------------------------------
def make_response_data(n, p, seed, cov=0):
    rng = np.random.default_rng(seed)
    c = uniform_coords(n, seed)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))])
    D = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    C = 5.0 * np.exp(-6.0 * D)
    w = np.linalg.cholesky(C + 1e-10 * np.eye(n)) @ rng.normal(size=n)
    beta = np.array([1.0, 5.0] + [0.5] * (p - 2))[:p]
    y = X @ beta + w + rng.normal(size=n)
    return c, X, y


# Exponential (a short and a longer chain), gaussian and matern; the tape
# positions check that the same number of normals and uniforms were drawn.
@pytest.mark.parametrize("cov,nu,S", [(0, 0.0, 200), (0, 0.0, 600),
                                      (3, 0.0, 300), (2, 1.0, 200)])
def test_rNNGP_chain_identical_to_oracle(oracle, nngp, cov, nu, S):
    n, m, p = 400, 10, 2
    print(f"rNNGP chain cov={cov} S={S}", flush=True)
    c, X, y = make_response_data(n, p, 42)
    nn = oracle.mkNNIndx(c, m)
    rng = np.random.default_rng(99)
    normals = rng.normal(size=S * 10 + 100)
    unifs = rng.uniform(size=S + 10)
    beta0 = np.linalg.lstsq(X, y, rcond=None)[0]
    args = dict(sigmaSqIG=[2.0, 5.0], tauSqIG=[2.0, 1.0], phiUnif=[3.0, 30.0],
                nuUnif=[0.5, 2.5], betaStarting=beta0, sigmaSqStarting=5.0,
                tauSqStarting=1.0, phiStarting=6.0, nuStarting=nu,
                sigmaSqTuning=0.3, tauSqTuning=0.3, phiTuning=0.3,
                nuTuning=0.2, nSamples=S)
    flat = c.T.reshape(-1).copy()
    ref_tape = oracle.TapeRng(normals, unifs)
    ref = oracle.rNNGP(y, X, p, n, m, flat, cov, nn["nnIndx"],
                       nn["nnIndxLU"], rng=ref_tape, **args)
    tape = oracle.TapeRng(normals, unifs)
    got = nngp.rNNGP(y, X, p, n, m, c, cov, nn["nnIndx"], nn["nnIndxLU"],
                     rng=tape, **args)
    # same number of draws consumed, same accept count, same chain
    assert (tape.i_n, tape.i_u) == (ref_tape.i_n, ref_tape.i_u)
    assert got["accept"] == ref["accept"]
    assert 0 < ref["accept"] < S
    th0, th1 = ref["p.theta.samples"], got["p.theta.samples"]
    acc0 = np.r_[True, np.any(th0[:, 1:] != th0[:, :-1], axis=0)]
    acc1 = np.r_[True, np.any(th1[:, 1:] != th1[:, :-1], axis=0)]
    assert np.array_equal(acc0, acc1)
    assert rel_err(th1, th0) < 1e-9
    assert rel_norm_err(got["p.beta.samples"], ref["p.beta.samples"]) < 1e-9


def test_rNNGP_verbose_reports(oracle, nngp, capfd):
    """The chain prints the reference's progress reports
    (src/rNNGP.cpp:360-373): at s = 100, 200 of 300, the interval rate
    counting iterations since the last report (the first interval has 101,
    as in the reference), then the final lines; every rate agrees with the
    accepted moves visible in the theta samples."""
    n, m, p, S = 300, 8, 2, 300
    c, X, y = make_response_data(n, p, 5)
    nn = oracle.mkNNIndx(c, m)
    beta0 = np.linalg.lstsq(X, y, rcond=None)[0]
    got = nngp.rNNGP(y, X, p, n, m, c, 0, nn["nnIndx"], nn["nnIndxLU"],
                     sigmaSqIG=[2.0, 5.0], tauSqIG=[2.0, 1.0],
                     phiUnif=[3.0, 30.0], nuUnif=None, betaStarting=beta0,
                     sigmaSqStarting=5.0, tauSqStarting=1.0, phiStarting=6.0,
                     nuStarting=0.0, sigmaSqTuning=0.3, tauSqTuning=0.3,
                     phiTuning=0.3, nuTuning=0.0, nSamples=S, verbose=1,
                     nReport=100, seed=3)
    out = capfd.readouterr().out
    th = got["p.theta.samples"]
    prev = np.r_[[5.0, 1.0, 6.0]][:, None]
    moved = np.any(th != np.c_[prev, th[:, :-1]], axis=0)
    cum = np.cumsum(moved)
    assert cum[-1] == got["accept"]
    want = []
    last = 0
    for s in (100, 200):
        want += ["Sampled: %i of %i, %3.2f%%" % (s, S, 100.0 * s / S),
                 "Report interval Metrop. Acceptance rate: %3.2f%%"
                 % (100.0 * (cum[s] - last) / 100),
                 "Overall Metrop. Acceptance rate: %3.2f%%"
                 % (100.0 * cum[s] / s),
                 "-------------------------------------------------"]
        last = cum[s]
    want += ["Sampled: %i of %i, %3.2f%%" % (S, S, 100.0),
             "Overall Metrop. Acceptance rate: %3.2f%%"
             % (100.0 * got["accept"] / S)]
    lines = out.splitlines()
    start = lines.index(want[0])
    assert lines[start:start + len(want)] == want


def test_rNNGP_fit_rep_matches_oracle(oracle, nngp):
    n, m, p, S = 200, 8, 2, 30
    c, X, y = make_response_data(n, p, 7)
    nn = oracle.mkNNIndx(c, m)
    rng = np.random.default_rng(3)
    normals = rng.normal(size=S * (n + 10) + 10)
    unifs = rng.uniform(size=S + 5)
    rep = np.zeros(S, np.int32)
    rep[10::5] = 1
    nrep = int(rep.sum())
    args = dict(sigmaSqIG=[2.0, 5.0], tauSqIG=[2.0, 1.0], phiUnif=[3.0, 30.0],
                nuUnif=None, betaStarting=np.zeros(p), sigmaSqStarting=5.0,
                tauSqStarting=1.0, phiStarting=6.0, nuStarting=0.0,
                sigmaSqTuning=0.3, tauSqTuning=0.3, phiTuning=0.3,
                nuTuning=0.0, nSamples=S, nRep=nrep, repIndx=rep)
    ref = oracle.rNNGP(y, X, p, n, m, c.T.reshape(-1).copy(), 0,
                       nn["nnIndx"], nn["nnIndxLU"],
                       rng=oracle.TapeRng(normals, unifs), **args)
    got = nngp.rNNGP(y, X, p, n, m, c, 0, nn["nnIndx"], nn["nnIndxLU"],
                     rng=oracle.TapeRng(normals, unifs), **args)
    assert got["accept"] == ref["accept"]
    assert rel_norm_err(got["y.rep.samples"], ref["y.rep.samples"]) < 1e-9


def test_rNNGP_fit_rep_at_1e5_matches_oracle(oracle, nngp):
    """fit.rep inside the chain at n = 1e5: the replicate draws run on the
    device from the chain's B, F and the host stream's normals
    (src/rNNGP.cpp:299-313); equal to the oracle's to 1e-10."""
    print("rNNGP fit.rep n=1e5", flush=True)
    n, m, p, S = 100_000, 15, 2, 4
    c = uniform_coords(n, 17)
    rng = np.random.default_rng(17)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, 2.0] + np.sin(7 * c[:, 0]) * np.cos(5 * c[:, 1]) \
        + 0.3 * rng.normal(size=n)
    nn = oracle.mkNNIndxCB(c, m, nThreads=16)
    normals = rng.normal(size=S * (n + 10) + 10)
    unifs = rng.uniform(size=S + 5)
    rep = np.array([0, 1, 0, 1], np.int32)
    args = dict(sigmaSqIG=[2.0, 1.0], tauSqIG=[2.0, 0.1], phiUnif=[3.0, 30.0],
                nuUnif=None, betaStarting=[1.0, 2.0], sigmaSqStarting=1.0,
                tauSqStarting=0.1, phiStarting=8.0, nuStarting=0.0,
                sigmaSqTuning=0.05, tauSqTuning=0.05, phiTuning=0.05,
                nuTuning=0.0, nSamples=S, nRep=int(rep.sum()), repIndx=rep)
    ref = oracle.rNNGP(y, X, p, n, m, c.T.reshape(-1).copy(), 0,
                       nn["nnIndx"], nn["nnIndxLU"], nThreads=16,
                       rng=oracle.TapeRng(normals, unifs), **args)
    got = nngp.rNNGP(y, X, p, n, m, c, 0, nn["nnIndx"], nn["nnIndxLU"],
                     rng=oracle.TapeRng(normals, unifs), **args)
    assert got["accept"] == ref["accept"]
    assert got["y.rep.samples"].shape == ref["y.rep.samples"].shape
    err = rel_norm_err(got["y.rep.samples"], ref["y.rep.samples"])
    print(f"  y.rep rel err {err:.1e}", flush=True)
    assert err < 1e-10


# ---------------------------------------------------------------------------
# kriging and the conjugate model
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("cov,m", [(0, 15), (0, 10), (3, 8), (1, 12),
                                   (2, 15), (0, 20), (0, 40)])
def test_rNNGPPredict_matches_oracle(oracle, nngp, cov, m):
    n, p, q, S = 3000, 3, 300, 70
    print(f"predict cov={cov} m={m}", flush=True)
    rng = np.random.default_rng(m + cov)
    c = rng.uniform(size=(n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))])
    y = X @ [1.0, 2.0, -1.0] + rng.normal(size=n)
    c0 = rng.uniform(size=(q, 2))
    X0 = np.column_stack([np.ones(q), rng.normal(size=(q, p - 1))])
    nn0, _ = oracle.nn2(c, c0, m)
    nTheta = 4 if cov == 2 else 3
    beta = rng.normal([1.0, 2.0, -1.0], 0.05, size=(S, p)).T.copy()
    th = np.vstack([np.exp(rng.normal(np.log(2.0), 0.1, S)),
                    np.exp(rng.normal(np.log(0.2), 0.1, S)),
                    np.exp(rng.normal(np.log(10.0), 0.1, S))]
                   + ([rng.uniform(0.6, 2.2, S)] if cov == 2 else []))
    z = rng.normal(size=q * S)
    ref = oracle.rNNGPPredict(X.T.reshape(-1).copy(), y,
                              c.T.reshape(-1).copy(), n, p, m,
                              X0.T.reshape(-1).copy(),
                              c0.T.reshape(-1).copy(), q,
                              nn0.T.reshape(-1).copy(),
                              beta.T.reshape(-1).copy(),
                              th.T.reshape(-1).copy(), S, cov, z)
    got = nngp.rNNGPPredict(X, y, c, n, p, m, X0, c0, q, nn0, beta, th, S,
                            cov, z=z)["p.y.0"]
    assert got.shape == (q, S)
    assert rel_norm_err(got, ref) < RTOL


@pytest.mark.parametrize("cov,nu", [(0, 0.0), (2, 1.5)])
def test_cNNGP_matches_oracle(oracle, nngp, cov, nu):
    n, m, p, n0, g = 2000, 15, 3, 400, 4
    rng = np.random.default_rng(17)
    c = uniform_coords(n, 17)
    X = rng.normal(size=(n, p))
    y = X @ [1.0, 2.0, -1.0] + rng.normal(size=n)
    c0 = rng.uniform(size=(n0, 2))
    X0 = rng.normal(size=(n0, p))
    nn = oracle.mkNNIndx(c, m)
    nn0, _ = oracle.nn2(c, c0, m)
    grid = [(3.0, 0.05), (8.0, 0.1), (15.0, 0.5), (30.0, 1.0)]
    ta = np.array([[ph, al] + ([nu] if cov == 2 else []) for ph, al in grid]).T
    ref = oracle.cNNGP(y, X.T.reshape(-1).copy(), p, n,
                       c.T.reshape(-1).copy(), ta.T.reshape(-1).copy(),
                       nn["nnIndx"], nn["nnIndxLU"], X0.T.reshape(-1).copy(),
                       c0.T.reshape(-1).copy(), n0,
                       nn0.T.reshape(-1).copy(), g, m, [2.0, 1.0], cov)
    got = nngp.cNNGP(y, X, p, n, c, ta, nn["nnIndx"], nn["nnIndxLU"], X0, c0,
                     n0, nn0, g, m, [2.0, 1.0], cov)
    for key in ("beta.hat", "ab", "bb", "y.0.hat", "y.0.hat.var"):
        assert rel_norm_err(got[key], ref[key]) < RTOL, key
    lower = np.tril_indices(p)
    for k in range(g):
        a = got["bB.inv"][:, k].reshape(p, p).T
        b = ref["bB.inv"][:, k].reshape(p, p).T
        assert rel_norm_err(a[lower], b[lower]) < RTOL


# ---------------------------------------------------------------------------
# error paths (forced from the host; every kernel still exits normally)
# ---------------------------------------------------------------------------
def test_not_positive_definite_is_reported(nngp, oracle):
    # duplicated coordinates with tau.sq = 0 make a singular neighbor system
    c = uniform_coords(300, 3)
    c[150] = c[149]
    nn = oracle.mkNNIndx(c, 5)
    with pytest.raises(nngp.NngpError) as e:
        nngp.updateBF(c, nn["nnIndx"], nn["nnIndxLU"], 5, [1.0, 0.0, 2.0], 0)
    assert e.value.status == nngp.NNGP_ERR_NOT_PD
    assert "dpotrf failed" in str(e.value)


def test_bad_cov_model_is_reported(nngp, oracle):
    c = uniform_coords(100, 3)
    nn = oracle.mkNNIndx(c, 5)
    with pytest.raises(nngp.NngpError) as e:
        nngp.updateBF(c, nn["nnIndx"], nn["nnIndxLU"], 5, [1.0, 0.1, 2.0], 9)
    assert e.value.status == nngp.NNGP_ERR_COV_MODEL
    assert "cov.model is not correctly specified" in str(e.value)


# ---------------------------------------------------------------------------
# mixed precision (FP32 build/factor/solve + one FP64 refinement step)
# ---------------------------------------------------------------------------
# The mixed form refines against its FP32-built covariance, so what remains
# against the FP64 oracle is the FP32 rounding of the covariance entries,
# amplified by the conditioning of the neighbor systems: the bounds below
# are those of that analysis (entries to ~1e-7, kappa ~ 1e2).
@pytest.mark.gpu
@pytest.mark.parametrize("cov,m", [(0, 15), (0, 8), (3, 10), (1, 12)])
def test_mixed_updateBF_close_to_oracle(oracle, nngp, cov, m):
    print(f"mixed updateBF cov={cov} m={m}", flush=True)
    n, p = 4000, 2
    c = uniform_coords(n, 40 + m)
    rng = np.random.default_rng(m)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, -1.0] + rng.normal(size=n)
    nn = oracle.mkNNIndxCB(c, m)
    theta = [2.0, 0.2, 12.0 if cov != 1 else 3.0]
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, theta, cov)
    beta = np.array([1.0, -1.0])
    r = X @ beta - y
    q0 = oracle.Q(B0, F0, r, r, nn["nnIndx"], nn["nnIndxLU"])
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, cov, X, y,
                    precision=1) as md:
        out = md.update(theta, slot=0, beta=beta)
        B, F = md.get_bf(0)
        # the same pass after tuning (any launch shape) agrees to rounding
        md.autotune(theta)
        out2 = md.update(theta, slot=1, beta=beta)
    ll0 = -0.5 * ld0 - 0.5 * q0
    ll = -0.5 * out["logDet"] - 0.5 * out["qResid"]
    assert abs(ll - ll0) / abs(ll0) < 1e-6
    assert np.max(np.abs(F - F0) / F0) < 1e-5
    assert np.max(np.abs(B - B0)) / np.max(np.abs(B0)) < 1e-4
    assert abs(out2["logDet"] - out["logDet"]) < 1e-9 * abs(out["logDet"])


@pytest.mark.gpu
@pytest.mark.parametrize("cov,m,p", [(0, 15, 3), (0, 15, 2), (3, 10, 1),
                                     (1, 12, 3), (0, 8, 3), (0, 15, 0),
                                     (2, 10, 3), (2, 16, 2)])
def test_mixed_lane_modes_agree(oracle, nngp, cov, m, p):
    """The mixed per-lane kernel's three ways to keep C for the refinement
    residual (re-evaluated, LDS copy, register copy at one wave per SIMD
    with the covariate rows staged in LDS where they can be) run the same
    arithmetic: B and F equal bit for bit across modes and block sizes, and
    the sums, added in another order by the register form, agree to
    rounding.  n = 4000 + 13 leaves a ragged last block; p = 2 has rows
    that stage in one 16-byte chunk, p = 1 and 0 none; at 256 threads the
    register form's stage of p = 3 rows exceeds 64 KiB and it reads them
    from memory.  The matern cases run the closed forms at nu = 1.5 and
    2.5."""
    print(f"mixed modes cov={cov} m={m} p={p}", flush=True)
    n = 4000 + 13
    c = uniform_coords(n, 60 + m + p)
    rng = np.random.default_rng(m + p)
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))]) \
        if p else None
    beta = np.linspace(1.0, -1.0, p) if p else np.zeros(1)
    y = (X @ beta if p else 0.0) + rng.normal(size=n)
    nn = oracle.mkNNIndxCB(c, m)
    theta = [2.0, 0.2, 12.0 if cov != 1 else 3.0]
    if cov == 2:
        theta.append(1.5 if m == 10 else 2.5)
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, theta, cov)
    ref = None
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, cov, X, y,
                    precision=1) as md:
        # the LDS copy (mode 1) fits a 64-thread block at m <= 16
        for th, mode in ((64, 0), (64, 1), (64, 3), (128, 0), (128, 3),
                         (256, 0), (256, 3)):
            md.set_shape(0, th, mode)
            out = md.update(theta, slot=0, beta=beta)
            B, F = md.get_bf(0)
            if ref is None:
                ref = (out, B, F)
                continue
            assert np.array_equal(B, ref[1]), (th, mode)
            assert np.array_equal(F, ref[2]), (th, mode)
            assert abs(out["logDet"] - ref[0]["logDet"]) \
                < 1e-12 * abs(ref[0]["logDet"])
            if p:
                assert abs(out["qResid"] - ref[0]["qResid"]) \
                    < 1e-12 * abs(ref[0]["qResid"])
    assert np.max(np.abs(ref[2] - F0) / F0) < 1e-5
    assert np.max(np.abs(ref[1] - B0)) / np.max(np.abs(B0)) < 1e-4
    assert abs(ref[0]["logDet"] - ld0) < 1e-6 * abs(ld0)


# The wavefront form (m > 16, or the Bessel-K matern at any m): FP32
# covariance build, including an FP32 Bessel K (Temme / Steed to 6e-8, a
# few ulp per entry), FP32 factor with one row per lane, one FP64 residual.
# Entries carry ~1e-6 relative error, the systems kappa ~ 1e2 (1e3 for the
# clustered sites), hence the bounds.
@pytest.mark.gpu
@pytest.mark.parametrize("cov,m,theta,clustered", [
    (0, 32, [2.0, 0.2, 12.0], False),
    (1, 32, [2.0, 0.2, 3.0], False),
    (3, 20, [1.0, 0.3, 5.0], False),
    (2, 15, [2.0, 0.2, 12.0, 1.2], False),
    (2, 20, [1.0, 0.1, 8.0, 0.7], False),
    (2, 32, [2.0, 0.2, 12.0, 2.5], False),
    (2, 32, [1.0, 0.05, 20.0, 2.5], True),
    (0, 48, [1.0, 0.1, 5.0], False),
])
def test_mixed_wavefront_close_to_oracle(oracle, nngp, cov, m, theta,
                                         clustered):
    print(f"mixed wavefront cov={cov} m={m} clustered={clustered}",
          flush=True)
    n = 4000
    rng = np.random.default_rng(40 + m + cov)
    if clustered:  # 50 tight clusters plus 10% uniform, x stretched by 3
        nc = int(0.9 * n)
        ce = rng.uniform(size=(50, 2))
        c = np.vstack([ce[rng.integers(0, 50, nc)]
                       + rng.normal(scale=0.02, size=(nc, 2)),
                       rng.uniform(size=(n - nc, 2))])
        c[:, 0] *= 3.0
    else:
        c = rng.uniform(size=(n, 2))
    c = c[np.argsort(c[:, 0], kind="stable")]
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, -1.0] + rng.normal(size=n)
    nn = oracle.mkNNIndxCB(c, m)
    B0, F0, ld0 = oracle.updateBF(c.T.reshape(-1).copy(), nn["nnIndx"],
                                  nn["nnIndxLU"], n, m, theta, cov)
    beta = np.array([1.0, -1.0])
    r = X @ beta - y
    q0 = oracle.Q(B0, F0, r, r, nn["nnIndx"], nn["nnIndxLU"])
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, cov, X, y,
                    precision=1) as md:
        out = md.update(theta, slot=0, beta=beta)
        B, F = md.get_bf(0)
        md.autotune(theta)
        out2 = md.update(theta, slot=1, beta=beta)
    ll0 = -0.5 * ld0 - 0.5 * q0
    ll = -0.5 * out["logDet"] - 0.5 * out["qResid"]
    assert abs(ll - ll0) / abs(ll0) < 1e-5
    assert np.max(np.abs(F - F0) / F0) < 5e-5
    assert np.max(np.abs(B - B0)) / np.max(np.abs(B0)) < 1e-4
    assert abs(out2["logDet"] - out["logDet"]) < 1e-9 * abs(out["logDet"])


@pytest.mark.gpu
def test_mixed_chain_matches_fp64_chain(oracle, nngp):
    """Same RNG tape: the mixed chain's posterior summaries equal the FP64
    chain's far inside Monte Carlo error."""
    print("mixed chain", flush=True)
    n, m, S = 3000, 10, 300
    c = uniform_coords(n, 77)
    rng = np.random.default_rng(77)
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ [1.0, 5.0] + np.sin(5 * c[:, 0]) + rng.normal(size=n)
    nn = oracle.mkNNIndxCB(c, m)
    args = dict(sigmaSqIG=[2, 1], tauSqIG=[2, 1], phiUnif=[3, 30],
                nuUnif=None, betaStarting=[1.0, 5.0], sigmaSqStarting=1.0,
                tauSqStarting=1.0, phiStarting=6.0, nuStarting=0.0,
                sigmaSqTuning=0.2, tauSqTuning=0.2, phiTuning=0.2,
                nuTuning=0.0, nSamples=S, seed=5)
    res = {}
    for prec in (0, 1):
        with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, 0, X, y,
                        precision=prec) as md:
            res[prec] = md.rNNGP(**args)
    for key in ("p.beta.samples", "p.theta.samples"):
        a, b = res[0][key], res[1][key]
        se = a.std(1) / np.sqrt(S)
        assert np.all(np.abs(a.mean(1) - b.mean(1)) < 0.01 * se + 1e-9)


@pytest.mark.gpu
def test_loop_timing_covers_the_iterations_only(oracle, nngp):
    """The loop clock counts each sampler call's iterations, not its initial
    pass; the event counters count every launch; enabling resets both."""
    print("loop timing", flush=True)
    n, m, p, S = 2000, 10, 2, 40
    c, X, y = make_response_data(n, p, 31)
    nn = oracle.mkNNIndxCB(c, m)
    args = dict(sigmaSqIG=[2.0, 5.0], tauSqIG=[2.0, 1.0], phiUnif=[3.0, 30.0],
                nuUnif=None, betaStarting=np.zeros(p), sigmaSqStarting=5.0,
                tauSqStarting=1.0, phiStarting=6.0, nuStarting=0.0,
                sigmaSqTuning=0.3, tauSqTuning=0.3, phiTuning=0.3,
                nuTuning=0.0, nSamples=S, seed=4)
    with nngp.Model(c, nn["nnIndx"], nn["nnIndxLU"], m, 0, X, y) as md:
        md.rNNGP(**args)  # untimed
        md.profile(True)
        t0 = time.perf_counter()
        md.rNNGP(**args)
        md.rNNGP(**args)
        wall_ms = (time.perf_counter() - t0) * 1e3
        tm = md.timing()
        assert tm["loop_iters"] == 2 * S
        assert tm["bf_launches"] == 2 * (S + 1)  # the initial pass each call
        assert tm["sum_launches"] == 2 * (S + 1)
        assert 0 < tm["loop_ms"] < wall_ms
        # every iteration holds its updateBF and sum kernels
        assert tm["loop_ms"] > 2 * S * tm["bf_ms"] / tm["bf_launches"]
        md.profile(True)
        tm = md.timing()
        assert tm["loop_iters"] == 0 and tm["loop_ms"] == 0.0
        assert tm["bf_launches"] == 0
        # sampled: every 5th updateBF launch and the sum launch after it
        md.profile(True, every=5)
        md.rNNGP(**args)
        md.rNNGP(**args)
        tm = md.timing()
        assert tm["loop_iters"] == 2 * S
        assert tm["bf_launches"] == 2 * (S + 1) // 5
        assert tm["sum_launches"] == tm["bf_launches"]
        assert tm["bf_ms"] > 0 and tm["sum_ms"] > 0


# ---------------------------------------------------------------------------
# interrupts (R_CheckUserInterrupt, src/rNNGP.cpp:377, src/cNNGP.cpp:314)
# ---------------------------------------------------------------------------
def test_interrupt_stops_the_chain_cleanly(oracle, nngp):
    n, m, p, S = 300, 8, 2, 50
    c, X, y = make_response_data(n, p, 21)
    nn = oracle.mkNNIndx(c, m)
    calls = []

    def stop_after_5():
        calls.append(1)
        return len(calls) >= 5

    args = dict(sigmaSqIG=[2.0, 5.0], tauSqIG=[2.0, 1.0], phiUnif=[3.0, 30.0],
                nuUnif=None, betaStarting=np.zeros(p), sigmaSqStarting=5.0,
                tauSqStarting=1.0, phiStarting=6.0, nuStarting=0.0,
                sigmaSqTuning=0.3, tauSqTuning=0.3, phiTuning=0.3,
                nuTuning=0.0, nSamples=S, seed=4)
    with pytest.raises(nngp.NngpError) as e:
        nngp.rNNGP(y, X, p, n, m, c, 0, nn["nnIndx"], nn["nnIndxLU"],
                   interrupt=stop_after_5, **args)
    assert e.value.status == nngp.NNGP_ERR_INTERRUPT
    assert "interrupted after 5 samples" in str(e.value)
    assert len(calls) == 5
    # the library is usable afterwards: the same chain runs to the end
    out = nngp.rNNGP(y, X, p, n, m, c, 0, nn["nnIndx"], nn["nnIndxLU"],
                     interrupt=lambda: False, **args)
    assert out["p.theta.samples"].shape == (3, S)


def test_interrupt_stops_the_grid_cleanly(oracle, nngp):
    n, m, p, q = 400, 6, 2, 30
    c, X, y = make_response_data(n, p, 22)
    nn = oracle.mkNNIndx(c, m)
    rng = np.random.default_rng(22)
    c0 = rng.uniform(size=(q, 2))
    X0 = np.column_stack([np.ones(q), rng.normal(size=q)])
    nn0, _ = oracle.nn2(c, c0, m)
    ta = np.array([[ph, al] for ph in (2.0, 5.0, 9.0) for al in (0.1, 0.5)]).T
    seen = []
    with pytest.raises(nngp.NngpError) as e:
        nngp.cNNGP(y, X, p, n, c, ta, nn["nnIndx"], nn["nnIndxLU"], X0, c0,
                   q, nn0, ta.shape[1], m, [2.0, 1.0], 0,
                   interrupt=lambda: seen.append(1) or len(seen) >= 2)
    assert e.value.status == nngp.NNGP_ERR_INTERRUPT
    assert "interrupted after 2 of 6" in str(e.value)
