"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/nngp.h declares, and rejects bad arguments before touching a GPU."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "nngp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nngp_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_what_the_binding_lists(nngp):
    from spnngp_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol(nngp):
    assert sorted(nngp.exported_symbols()) == declared_symbols()


def test_version_and_device_count(nngp):
    assert nngp.load().nngp_version() == 10000
    assert nngp.device_count() >= 0


def test_bad_arguments_rejected_without_gpu(nngp):
    c = np.random.default_rng(0).uniform(size=(50, 2))
    nnIndx = np.zeros(nngp.n_indx(50, 5), np.int32)
    lu = np.zeros(100, np.int32)
    # malformed nnIndxLU (count above m) is an argument error
    lu[50 + 3] = 7
    with pytest.raises(nngp.NngpError) as e:
        nngp.updateBF(c, nnIndx, lu, 5, [1.0, 0.1, 1.0], 0)
    assert e.value.status == nngp.NNGP_ERR_ARG
    with pytest.raises(nngp.NngpError) as e:
        nngp.updateBF(c, nnIndx, lu, 5, [1.0, 0.1, 1.0], 5)
    assert e.value.status == nngp.NNGP_ERR_COV_MODEL
    with pytest.raises(nngp.NngpError) as e:
        nngp.mkNNIndxCB(c, 50)  # m must be below n
    assert e.value.status == nngp.NNGP_ERR_ARG


def test_no_device_is_a_clean_error(nngp):
    if nngp.device_count() > 0:
        pytest.skip("a device is visible")
    c = np.random.default_rng(0).uniform(size=(50, 2))
    with pytest.raises(nngp.NngpError) as e:
        nngp.mkNNIndxCB(c, 5)
    assert e.value.status == nngp.NNGP_ERR_HIP


def test_sNNGP_rejects_a_bad_transpose_before_the_device(nngp, oracle):
    n, m = 100, 5
    rng = np.random.default_rng(3)
    c = rng.random((n, 2))
    c = c[np.argsort(c[:, 0], kind="stable")]
    r = oracle.mkNNIndxCB(c, m)
    U = oracle.mkUIndx(n, m, r["nnIndx"], r["nnIndxLU"], 2)
    bad = U["ui.indx"].copy()
    bad[7] = (bad[7] + 1) % 5  # now names a slot holding another site
    X = np.ones((n, 1))
    with pytest.raises(nngp.NngpError) as e:
        nngp.sNNGP(np.zeros(n), X, 1, n, m, c, 0, r["nnIndx"], r["nnIndxLU"],
                   U["u.indx"], U["u.indx.lu"], bad, [2, 1], [2, 1], [1, 30],
                   None, [0.0], 1, 1, 5, 0, 0, 0, 0.1, 0, 2)
    assert e.value.status == nngp.NNGP_ERR_ARG
    with pytest.raises(nngp.NngpError) as e:  # unknown sweep order
        nngp.sNNGP(np.zeros(n), X, 1, n, m, c, 0, r["nnIndx"], r["nnIndxLU"],
                   U["u.indx"], U["u.indx.lu"], U["ui.indx"], [2, 1], [2, 1],
                   [1, 30], None, [0.0], 1, 1, 5, 0, 0, 0, 0.1, 0, 2, sweep=7)
    assert e.value.status == nngp.NNGP_ERR_ARG


def test_latent_entry_points_reject_bad_arguments(nngp):
    import ctypes as C
    L = nngp.load()
    nn = np.zeros(1, np.int32)
    lu = np.zeros(4, np.int32)
    u = np.zeros(1, np.int32)
    # n < 2 and m >= n are argument errors, before any device call
    assert L.nngp_mkUIndx(1, 1, nn, u, lu, u, lu, 2) == nngp.NNGP_ERR_ARG
    assert L.nngp_mkUIndx(2, 2, nn, u, lu, u, lu, 2) == nngp.NNGP_ERR_ARG
    with pytest.raises(nngp.NngpError) as e:
        nngp.sNNGP(np.zeros(3), np.ones((3, 1)), 1, 3, 1, np.zeros((3, 2)),
                   7, np.zeros(2, np.int32), np.zeros(6, np.int32),
                   np.zeros(2, np.int32), np.zeros(6, np.int32),
                   np.zeros(2, np.int32), [2, 1], [2, 1], [1, 30], None,
                   [0.0], 1, 1, 5, 0, 0, 0, 0.1, 0, 1)
    assert e.value.status == nngp.NNGP_ERR_COV_MODEL
    with pytest.raises(nngp.NngpError) as e:  # p = 0
        nngp.sNNGP(np.zeros(3), np.ones((3, 1)), 0, 3, 1, np.zeros((3, 2)),
                   0, np.zeros(2, np.int32), np.zeros(6, np.int32),
                   np.zeros(2, np.int32), np.zeros(6, np.int32),
                   np.zeros(2, np.int32), [2, 1], [2, 1], [1, 30], None,
                   [0.0], 1, 1, 5, 0, 0, 0, 0.1, 0, 1)
    assert e.value.status == nngp.NNGP_ERR_UNSUPPORTED
    with pytest.raises(nngp.NngpError) as e:  # binomial family
        nngp.sNNGPPredict(np.ones((3, 1)), None, np.zeros((3, 2)), 3, 1, 1,
                          np.ones((1, 1)), np.zeros((1, 2)), 1,
                          np.zeros((1, 1), np.int32), np.zeros((1, 1)),
                          np.ones((3, 1)), np.zeros((3, 1)), 1, 0, family=2)
    assert e.value.status == nngp.NNGP_ERR_UNSUPPORTED
