"""GPU neighbor search on inputs with exactly tied distances, bit for bit
against the oracle's restatement of each search: mkNNIndxCB against the
code-book walk (src/nn.cpp:136-257), mkNNIndx against the scan
(src/nn.cpp:63-111), nn2 against the scan ordered by (distance, row).

Regular lattices, coordinates rounded to a few decimals and duplicated
sites make most sites' lists depend on how ties are visited and sorted, so
these inputs exercise the replay of the reference's own candidate order.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def x_ordered(c):
    return c[np.argsort(c[:, 0], kind="stable")]


def lattice(k):
    g = np.arange(k) / k
    return x_ordered(np.array([(a, b) for a in g for b in g]))


def rounded(n, dec, seed):
    rng = np.random.default_rng(seed)
    return x_ordered(np.round(rng.uniform(size=(n, 2)), dec))


def duplicated(seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(size=(3000, 2))
    return x_ordered(np.vstack([c, c[:500], c[:50], c[:50]]))


INPUTS = {
    "lattice120": lambda: lattice(120),
    "rounded3": lambda: rounded(8000, 3, 31),
    "duplicates": lambda: duplicated(32),
}


def same(got, ref):
    assert np.array_equal(got["nnIndxLU"], ref["nnIndxLU"])
    bad = np.flatnonzero(got["nnIndx"] != ref["nnIndx"])
    assert bad.size == 0, f"{bad.size} index entries differ, first {bad[:5]}"
    assert np.array_equal(got["nnDist"], ref["nnDist"])  # bit for bit


@pytest.mark.parametrize("m", [10, 15, 32])
@pytest.mark.parametrize("kind", list(INPUTS))
def test_mkNNIndxCB_ties_bit_exact(oracle, nngp, kind, m):
    print(f"mkNNIndxCB ties {kind} m={m}", flush=True)
    c = INPUTS[kind]()
    same(nngp.mkNNIndxCB(c, m), oracle.mkNNIndxCB(c, m, nThreads=8))


@pytest.mark.parametrize("m", [10, 15, 32])
@pytest.mark.parametrize("kind", list(INPUTS))
def test_mkNNIndx_ties_bit_exact(oracle, nngp, kind, m):
    print(f"mkNNIndx ties {kind} m={m}", flush=True)
    c = INPUTS[kind]()
    same(nngp.mkNNIndx(c, m), oracle.mkNNIndx(c, m, nThreads=8))


def test_ties_in_a_custom_order(oracle, nngp):
    # NNGP order unrelated to x: predecessors by index, u ties everywhere
    rng = np.random.default_rng(33)
    c = lattice(60)[rng.permutation(3600)]
    for m in (8, 20):
        same(nngp.mkNNIndxCB(c, m), oracle.mkNNIndxCB(c, m, nThreads=8))
        same(nngp.mkNNIndx(c, m), oracle.mkNNIndx(c, m, nThreads=8))


def test_ties_known_answer_2x2(nngp):
    # hand-traced through src/nn.cpp: the walk lists s3's neighbors as
    # (s2, s1), the scan as (s1, s2)
    c = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0], [1.0, 1.0]])
    assert list(nngp.mkNNIndxCB(c, 2)["nnIndx"]) == [0, 0, 1, 2, 1]
    assert list(nngp.mkNNIndx(c, 2)["nnIndx"]) == [0, 0, 1, 1, 2]


def test_continuous_inputs_unchanged_by_ties_path(oracle, nngp):
    # no ties: the code-book and scan outputs coincide, as before
    rng = np.random.default_rng(34)
    c = x_ordered(rng.uniform(size=(6000, 2)))
    a = nngp.mkNNIndxCB(c, 15)
    same(a, oracle.mkNNIndxCB(c, 15, nThreads=8))
    same(nngp.mkNNIndx(c, 15), a)


@pytest.mark.parametrize("k", [1, 10, 15, 40])
def test_nn2_ties_bit_exact(oracle, nngp, k):
    print(f"nn2 ties k={k}", flush=True)
    c = lattice(60)[::-1].copy()
    rng = np.random.default_rng(35 + k)
    c0 = np.vstack([c[::13] + 1.0 / 120, c[::17],
                    np.round(rng.uniform(size=(300, 2)), 2)])
    i_ref, d_ref = oracle.nn2(c, c0, k)
    i_got, d_got = nngp.nn2(c, c0, k)
    assert np.array_equal(i_got, i_ref)
    assert np.array_equal(d_got, d_ref)
