"""R's default generator in the library (nngp_rng_r_create): known answers
of R's set.seed(seed); runif(k) / rnorm(k) as R prints them (7 significant
digits; RNGkind "Mersenne-Twister", normal.kind "Inversion"), and the
moments of its normals.  Host code only: nothing here touches a GPU."""
import numpy as np
import pytest

import spnngp_amd as nn

# R >= 3.6 output of set.seed(s); runif(k) or rnorm(k)
KAT = [
    (1, "unif", [0.2655087, 0.3721239, 0.5728534]),
    (1, "norm", [-0.6264538, 0.1836433, -0.8356286, 1.5952808, 0.3295078]),
    (42, "norm", [1.3709584, -0.5646982, 0.3631284]),
    (123, "norm", [-0.56047565, -0.23017749, 1.55870831]),
    (123, "unif", [0.2875775, 0.7883051, 0.4089769]),
]


@pytest.mark.parametrize("seed,kind,want", KAT)
def test_matches_r_set_seed_streams(seed, kind, want):
    r = nn.RRng(seed)
    got = r.unif(len(want)) if kind == "unif" else r.norm(len(want))
    # R prints at least 7 significant digits: within half a unit of the 7th
    want = np.asarray(want)
    half_ulp7 = 0.5 * 10.0 ** (np.floor(np.log10(np.abs(want))) - 6)
    assert np.all(np.abs(got - want) <= half_ulp7 * (1 + 1e-9))


def test_normals_have_standard_moments():
    z = nn.RRng(7).norm(200_000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    assert abs(np.mean(z ** 3)) < 0.03 and abs(np.mean(z ** 4) - 3) < 0.06


def test_streams_are_independent_objects():
    a, b = nn.RRng(99), nn.RRng(99)
    x = a.norm(3)
    assert np.array_equal(b.norm(3), x)   # same seed, same stream
    assert not np.array_equal(a.norm(3), x)
