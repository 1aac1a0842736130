"""nngp_Q_noalloc (device buffers in and out, partial sums in the library's
file-scope device array) against nngp_Q (host buffers, a scratch buffer
allocated per call) through the C ABI: two calls in a row, the second with
a smaller input, each bit for bit the existing entry point's result (both
run the same kernels over the same reduction grid)."""
import ctypes as C

import numpy as np
import pytest

from conftest import uniform_coords

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """Device buffers through the HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def case(oracle, nngp, n, m, seed):
    c = uniform_coords(n, seed)
    rng = np.random.default_rng(seed)
    nn = oracle.mkNNIndxCB(c, m, nThreads=16)
    B, F, _ = nngp.updateBF(c, nn["nnIndx"], nn["nnIndxLU"], m,
                            [2.0, 0.2, 12.0], "exponential")
    u, v = rng.normal(size=n), rng.normal(size=n)
    return B, F, u, v, nn


@pytest.mark.parametrize("sizes", [((200_000, 15), (3_000, 10))])
def test_Q_noalloc_matches_Q_twice(oracle, nngp, sizes):
    L = nngp.load()
    hip = Hip()
    try:
        for k, (n, m) in enumerate(sizes):
            print(f"Q_noalloc call {k}: n={n} m={m}", flush=True)
            B, F, u, v, nn = case(oracle, nngp, n, m, 60 + k)
            want = nngp.Q(B, F, u, v, nn["nnIndx"], nn["nnIndxLU"], m)
            dq = hip.put(np.full(1, np.nan))
            st = L.nngp_Q_noalloc(hip.put(B), hip.put(F), hip.put(u),
                                  hip.put(v), n, m, hip.put(nn["nnIndx"]),
                                  hip.put(nn["nnIndxLU"]), dq, None)
            assert st == 0, L.nngp_last_error()
            assert hip.L.hipDeviceSynchronize() == 0
            got = hip.get(dq, 1, np.float64)[0]
            assert got == want, (k, got, want)
            ref = oracle.Q(B, F, u, v, nn["nnIndx"], nn["nnIndxLU"])
            assert abs(got - ref) <= 1e-10 * abs(ref)
            hip.free()
    finally:
        hip.free()


def test_Q_noalloc_rejects_bad_arguments(nngp):
    L = nngp.load()
    st = L.nngp_Q_noalloc(None, None, None, None, 100, 5, None, None, None,
                          None)
    assert st == nngp.NNGP_ERR_ARG
    # nIndx = m(m+1)/2 + (n-m-1)m beyond int32
    one = C.c_void_p(8)
    st = L.nngp_Q_noalloc(one, one, one, one, 2_000_000_000, 60, one, one,
                          one, None)
    assert st == nngp.NNGP_ERR_ARG
    assert b"2^31" in L.nngp_last_error()
