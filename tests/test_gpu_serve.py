"""Kriging server (nngp_server_*): results against the oracle's
rNNGPPredict with the server's counter-based normals, bit-identical results
alone and in a mix of concurrent requests, and footprint admission."""
import threading

import numpy as np
import pytest

RTOL = 1e-10


def max_rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def fitted(cov, n=3000, p=3, S=40, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(size=(n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, p - 1))])
    y = X @ [1.0, 2.0, -1.0][:p] + rng.normal(size=n)
    beta = rng.normal([1.0, 2.0, -1.0][:p], 0.05, size=(S, p)).T.copy()
    th = np.vstack([np.exp(rng.normal(np.log(2.0), 0.1, S)),
                    np.exp(rng.normal(np.log(0.2), 0.1, S)),
                    np.exp(rng.normal(np.log(10.0), 0.1, S))]
                   + ([rng.uniform(0.6, 2.2, S)] if cov == 2 else []))
    return c, X, y, beta, th


def request(q, p, seed):
    rng = np.random.default_rng(1000 + seed)
    c0 = rng.uniform(size=(q, 2))
    X0 = np.column_stack([np.ones(q), rng.normal(size=(q, p - 1))])
    return X0, c0


@pytest.mark.gpu
@pytest.mark.parametrize("cov,m", [(0, 10), (0, 15), (1, 8), (3, 12),
                                   (2, 10), (0, 20)])
def test_server_matches_oracle(oracle, nngp, cov, m):
    print(f"server vs oracle cov={cov} m={m}", flush=True)
    n, p, S, q = 3000, 3, 40, 157
    c, X, y, beta, th = fitted(cov, n, p, S, seed=m + cov)
    X0, c0 = request(q, p, seed=cov * 100 + m)
    seed = 0x1234ABCD00000000 + 17 * m + cov
    with nngp.Server(X, y, c, m, beta, th, cov, device_bytes=64 << 20,
                     host_bytes=64 << 20) as srv:
        got = srv.predict(X0, c0, seed)["p.y.0"]
    nn0, _ = oracle.nn2(c, c0, m)
    z = oracle.philox_normals(seed, q, S)
    ref = oracle.rNNGPPredict(X.T.reshape(-1).copy(), y,
                              c.T.reshape(-1).copy(), n, p, m,
                              X0.T.reshape(-1).copy(),
                              c0.T.reshape(-1).copy(), q,
                              nn0.T.reshape(-1).copy(),
                              beta.T.reshape(-1).copy(),
                              th.T.reshape(-1).copy(), S, cov, z)
    assert got.shape == (q, S)
    assert max_rel(got, ref) < RTOL


@pytest.mark.gpu
def test_server_result_same_alone_and_in_mix(nngp):
    """A request's result is bit-identical when it runs alone and when it
    shares launches with many concurrent requests of uneven size."""
    print("server alone vs mix", flush=True)
    n, p, S, m = 4000, 3, 60, 15
    c, X, y, beta, th = fitted(0, n, p, S, seed=5)
    probe_q = 1234
    Xp, cp = request(probe_q, p, seed=77)
    with nngp.Server(X, y, c, m, beta, th, 0, device_bytes=96 << 20,
                     host_bytes=96 << 20, batch_sites=500, depth=3) as srv:
        alone = srv.predict(Xp, cp, 99)["p.y.0"]
        rng = np.random.default_rng(3)
        sizes = [int(v) for v in np.exp(rng.uniform(0, np.log(3000), 40))]
        results, errors = {}, []

        def client(k0):
            try:
                for k in range(k0, len(sizes), 3):
                    Xr, cr = request(sizes[k], p, seed=k)
                    t = srv.submit(Xr, cr, 1000 + k)
                    results[k] = t
                    if k == 19:  # the probe goes in the middle of the mix
                        results["probe"] = srv.submit(Xp, cp, 99)
            except Exception as e:  # pragma: no cover - reported below
                errors.append(e)

        ths = [threading.Thread(target=client, args=(k,)) for k in range(3)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        assert not errors, errors
        got = {}
        for _ in range(len(results)):
            t, info = srv.wait(-1, timeout=60.0)
            assert info["status"] == 0
            got[t] = info
            srv.release(t)
        st = srv.stats()
    mixed = got[results["probe"]]["y0"]
    assert np.array_equal(mixed, alone)
    assert st["completed"] == len(sizes) + 2
    assert st["launches"] > 5
    # every request of the mix finished with its own q
    for k, t in results.items():
        if k != "probe":
            assert got[t]["q"] == sizes[k]
            assert got[t]["t_done"] >= got[t]["t_admit"] >= got[t]["t_submit"]


@pytest.mark.gpu
def test_server_admits_by_footprint(nngp):
    """With a device arena that holds only a few requests at a time, a burst
    of requests is admitted as room frees up, and every one completes."""
    print("server admission", flush=True)
    n, p, S, m = 2000, 2, 50, 10
    c, X, y, beta, th = fitted(0, n, p, S, seed=9)
    q = 2000
    with nngp.Server(X, y, c, m, beta, th, 0, device_bytes=3 << 20,
                     host_bytes=64 << 20, batch_sites=700) as srv:
        per = 16 * q + 8 * p * q + 4 * m * q + 8 * q * S
        assert 2 * per <= 3 << 20 < 4 * per
        X0, c0 = request(q, p, seed=1)
        ref = srv.predict(X0, c0, 5)["p.y.0"]
        tickets = [srv.submit(X0, c0, 5) for _ in range(12)]
        for t in tickets:
            _, info = srv.wait(t, timeout=60.0)
            assert info["status"] == 0
            assert np.array_equal(info["y0"], ref)
            srv.release(t)
        st = srv.stats()
        assert st["device_bytes_peak"] <= 3 << 20
        assert st["device_bytes_used"] == 0
        # a request that can never fit is refused at submission
        Xb, cb = request(20000, p, seed=2)
        with pytest.raises(nngp.NngpError) as e:
            srv.submit(Xb, cb, 1)
        assert e.value.status == nngp.NNGP_ERR_NOMEM


@pytest.mark.gpu
def test_server_wait_timeout_and_unknown_ticket(nngp):
    print("server wait edge cases", flush=True)
    c, X, y, beta, th = fitted(0, 500, 2, 10, seed=11)
    with nngp.Server(X, y, c, 5, beta, th, 0, device_bytes=8 << 20,
                     host_bytes=8 << 20) as srv:
        with pytest.raises(nngp.NngpError) as e:
            srv.wait(-1, timeout=0.01)
        assert e.value.status == 8  # NNGP_ERR_TIMEOUT
        with pytest.raises(nngp.NngpError):
            srv.wait(12345, timeout=0.01)
        X0, c0 = request(3, 2, seed=4)
        t = srv.submit(X0, c0, 1)
        _, info = srv.wait(t, timeout=30.0)
        assert info["y0"].shape == (3, 10)
        srv.release(t)
        with pytest.raises(nngp.NngpError):
            srv.release(t)


@pytest.mark.gpu
def test_server_done_queue_stays_bounded(nngp):
    """Requests waited for by ticket (or served by predict) leave the queue
    of finished requests: after many of them it is empty, and a wait for
    any request then times out instead of returning a stale ticket."""
    print("server done queue", flush=True)
    c, X, y, beta, th = fitted(0, 800, 2, 8, seed=12)
    with nngp.Server(X, y, c, 6, beta, th, 0, device_bytes=16 << 20,
                     host_bytes=16 << 20) as srv:
        tickets = []
        for k in range(200):
            X0, c0 = request(1 + k % 7, 2, seed=100 + k)
            tickets.append(srv.submit(X0, c0, k))
        for t in tickets:
            srv.wait(t, timeout=30.0, copy=False)
            srv.release(t)
        for k in range(50):
            X0, c0 = request(3, 2, seed=400 + k)
            srv.predict(X0, c0, k)
        # finished, released before any wait: also leaves the queue
        X0, c0 = request(2, 2, seed=999)
        t = srv.submit(X0, c0, 9)
        st = srv.stats()
        while st["completed"] < 251:
            st = srv.stats()
        srv.release(t)
        st = srv.stats()
        assert st["done_unreported"] == 0
        assert st["completed"] == 251
        assert st["kernel_ms"] > 0.0
        assert st["kernel_site_samples"] == st["site_samples"]
        with pytest.raises(nngp.NngpError) as e:
            srv.wait(-1, timeout=0.05)
        assert e.value.status == 8  # NNGP_ERR_TIMEOUT
