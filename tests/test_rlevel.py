"""Host-side pieces of the R-level front ends (R/util.R, R/spNNGP.R):
the CRPS of R/util.R:247-251 against its definition as an integral, RMSPE,
sub-sampling and the 1-based neighbor lists of R/util.R:272-279."""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.stats as ss

from spnngp_amd import spnngp as R


def test_crps_is_the_integral_definition():
    # CRPS(N(mu, s^2), y) = int (F(x) - 1{x >= y})^2 dx
    for mu, var, y in [(0.0, 1.0, 0.3), (2.0, 0.25, -1.0), (-1.0, 4.0, 5.0)]:
        sd = np.sqrt(var)
        f = lambda x: (ss.norm.cdf(x, mu, sd) - (x >= y)) ** 2
        lo, hi = mu - 40 * sd, mu + 40 * sd
        want = (si.quad(f, lo, y, limit=200)[0]
                + si.quad(f, y, hi, limit=200)[0])
        assert R.crps([y], [mu], [var]) == pytest.approx(want, rel=1e-9)


def test_crps_and_rmspe_average_over_sites():
    rng = np.random.default_rng(0)
    y, yh, v = rng.normal(size=50), rng.normal(size=50), rng.uniform(1, 2, 50)
    each = [R.crps([a], [b], [c]) for a, b, c in zip(y, yh, v)]
    assert R.crps(y, yh, v) == pytest.approx(np.mean(each), rel=1e-13)
    assert R.rmspe(y, yh) == pytest.approx(
        np.sqrt(np.mean((y - yh) ** 2)), rel=1e-15)


def test_sub_sample_follows_R_seq():
    ss_, idx = R._sub_sample({"start": 3, "end": 11, "thin": 4}, 20, 1)
    assert list(idx) == [3, 7, 11]            # seq(3, 11, by = 4)
    _, idx = R._sub_sample(None, 10, 5)
    assert list(idx) == list(range(5, 11))
    with pytest.raises(ValueError):
        R._sub_sample({"start": 10}, 10, 1)


def test_n_indx_list_matches_layout(oracle):
    n, m = 30, 4
    c = np.random.default_rng(1).uniform(size=(n, 2))
    nn = oracle.mkNNIndx(c[np.argsort(c[:, 0])], m)
    lst = R.n_indx_list(nn["nnIndx"], nn["nnIndxLU"], n)
    assert lst[0] is None
    for i in range(1, n):
        off, cnt = oracle.getNNIndx(i, m)
        assert list(lst[i]) == list(nn["nnIndx"][off:off + cnt] + 1)


def _fake_fit(n=50, S=20, m=5, seed=0):
    rng = np.random.default_rng(seed)
    lu = np.zeros(2 * n, np.int32)
    idx = []
    for i in range(n):
        k = min(i, m)
        lu[i], lu[n + i] = len(idx), k
        idx += list(range(i - k, i))
    return {
        "p.beta.samples": rng.normal(size=(S, 2)),
        "p.theta.samples": rng.uniform(0.5, 2, size=(S, 3)),
        "X": rng.normal(size=(n, 2)), "y": rng.normal(size=n),
        "coords": rng.uniform(size=(n, 2)),
        "theta.names": ["sigma.sq", "tau.sq", "phi"],
        "x.names": ["x1", "x2"], "run.time": 1.5, "accept": 7,
        "n.neighbors": m, "cov.model": "exponential", "cov.model.indx": 0,
        "starting": {"sigma.sq": 1.0, "tau.sq": 0.5, "phi": 6.0},
        "type": ["response", "gaussian"],
        "neighbor.info": {"n.neighbors": m,
                          "nn.indx": np.array(idx, np.int32),
                          "nn.indx.lu": lu, "ord": np.arange(n),
                          "nn.indx.run.time": 0.01},
    }


def test_fit_file_round_trip_without_pickle(tmp_path):
    fit = _fake_fit()
    path = tmp_path / "fit.npz"
    R.save_fit(fit, path)
    # readable with pickle disabled: only numeric arrays and one string
    with np.load(path, allow_pickle=False) as f:
        assert all(f[k].dtype != object for k in f.files)
    back = R.load_fit(path)
    for k, v in fit.items():
        if k == "neighbor.info":
            for kk, vv in v.items():
                got = back[k][kk]
                assert np.array_equal(got, vv) if isinstance(vv, np.ndarray) \
                    else got == vv
        elif isinstance(v, np.ndarray):
            assert np.array_equal(back[k], v) and back[k].dtype == v.dtype
        else:
            assert back[k] == v
    # the 1-based neighbor lists of R/util.R:272-279 are rebuilt on load
    ni = back["neighbor.info"]
    want = R.n_indx_list(ni["nn.indx"], ni["nn.indx.lu"], 50)
    assert all((a is None and b is None) or np.array_equal(a, b)
               for a, b in zip(ni["n.indx"], want))


def test_fit_file_rejects_object_arrays(tmp_path):
    fit = _fake_fit()
    fit["bad"] = np.array([{"a": 1}], dtype=object)
    with pytest.raises(ValueError):
        R.save_fit(fit, tmp_path / "bad.npz")


def test_latent_fit_round_trip_and_its_guards(tmp_path):
    """A latent fit (p.w.samples, the neighbors-of-neighbors index, the
    sweep order) survives the fit file; serve() and a server for predict()
    are refused for it, before any device work."""
    fit = _fake_fit()
    rng = np.random.default_rng(5)
    n, S = fit["X"].shape[0], fit["p.beta.samples"].shape[0]
    fit["type"] = ["latent", "gaussian"]
    fit["sweep"] = "coloured"
    fit["p.w.samples"] = rng.normal(size=(n, S))
    fit["neighbor.info"]["u.indx"] = np.arange(10, dtype=np.int32)
    fit["neighbor.info"]["ui.indx"] = np.zeros(10, np.int32)
    path = tmp_path / "latent.npz"
    R.save_fit(fit, path)
    back = R.load_fit(path)
    assert back["type"] == ["latent", "gaussian"]
    assert back["sweep"] == "coloured"
    assert np.array_equal(back["p.w.samples"], fit["p.w.samples"])
    assert np.array_equal(back["neighbor.info"]["u.indx"],
                          fit["neighbor.info"]["u.indx"])
    with pytest.raises(ValueError):
        R.serve(back)
    with pytest.raises(ValueError):
        R.predict(back, np.ones((2, 2)), np.zeros((2, 2)), server=object())
