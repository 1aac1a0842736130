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
