"""numpy interface to libnngp_hip.so, named after spNNGP's native routines.

Each function takes the arguments of the reference routine it replaces, in
the reference's layouts (column-major FP64 matrices, int32 flat neighbor
index), and returns what the `.Call` returns, as numpy arrays.  All compute
runs in the HIP library; errors raise NngpError with the library's message.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import NngpOpts, NngpPassOut, NngpRng, check, load

COV_MODELS = ("exponential", "spherical", "matern", "gaussian")


def cov_index(cov) -> int:
    """covModel index of src/util.cpp:192-206 from a name or an int."""
    if isinstance(cov, str):
        return COV_MODELS.index(cov)
    return int(cov)


def n_indx(n: int, m: int) -> int:
    """Length of the flat neighbor arrays (src/rNNGP.cpp:218)."""
    return int((1 + m) / 2 * m + (n - m - 1) * m)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cm(a):
    """Column-major flat copy of a 2-D array (or a flat array unchanged)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 2:
        return _f64(a.T.reshape(-1))
    return _f64(a)


def _opts(precision: int = 0, device: int = 0, seed: int = 1, shape=None,
          sweep: int = 0, search: int = 0, interrupt=None):
    """nngp_opts.  interrupt: a callable polled once per MCMC iteration or
    grid point (nonzero stops the call with NNGP_ERR_INTERRUPT); the
    returned structure keeps its C wrapper alive."""
    sh = (C.c_int * 3)(*(shape if shape is not None else (0, 0, 0)))
    fn = _lib._INT_FN(lambda _arg: int(bool(interrupt()))) if interrupt \
        else _lib._INT_FN()
    o = NngpOpts(precision, device, seed, fn, None, sh, sweep, search,
                 (C.c_int * 3)())
    o._keep = fn
    return o


def device_count() -> int:
    return load().nngp_device_count()


# ---------------------------------------------------------------------------
# neighbor search
# ---------------------------------------------------------------------------
def _knn(fn, coords, m):
    coords = _cm(coords)
    n = coords.size // 2
    ni = n_indx(n, m)
    nnIndx = np.zeros(ni, np.int32)
    nnDist = np.zeros(ni, np.float64)
    nnIndxLU = np.zeros(2 * n, np.int32)
    check(fn(n, m, coords, nnIndx, nnDist, nnIndxLU, 1))
    return {"nnIndx": nnIndx, "nnDist": nnDist, "nnIndxLU": nnIndxLU}


def mkNNIndxCB(coords, m: int):
    """mkNNIndxCB (R/nn.R:53-76, src/nn.cpp:214-257).  coords: n x 2."""
    return _knn(load().nngp_mkNNIndxCB, coords, m)


def mkNNIndx(coords, m: int):
    """mkNNIndx (R/nn.R:27-51, src/nn.cpp:63-111)."""
    return _knn(load().nngp_mkNNIndx, coords, m)


def nn2(coords, coords0, k: int):
    """RANN::nn2(coords, coords0, k) with 0-based indices (q x k)."""
    coords = np.asarray(coords, np.float64)
    coords0 = np.asarray(coords0, np.float64)
    n, q = coords.shape[0], coords0.shape[0]
    idx = np.zeros(q * k, np.int32)
    dist = np.zeros(q * k, np.float64)
    check(load().nngp_nn2(n, _cm(coords), q, _cm(coords0), k, idx,
                          dist.ctypes.data_as(C.c_void_p)))
    return idx.reshape(k, q).T.copy(), dist.reshape(k, q).T.copy()


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------
def _shape_name(family, mode):
    """Name of an updateBF launch shape of the autotuner: 'lane' (one site
    per lane; '+lds-copy' and '+reg-copy' for the mixed form's covariance
    copy in LDS or in registers, '+lds-z' for the FP64 form's covariate rows
    staged in LDS), 'wave' (one site per wavefront) or 'row' (one site per
    16-lane row, four per wavefront)."""
    if family == 0:
        return ("lane", "lane+lds-copy", "lane+lds-z", "lane+reg-copy")[mode]
    return "row" if family == 2 else "wave"


def updateBF(coords, nnIndx, nnIndxLU, m: int, theta, covModel,
             precision: int = 0):
    """updateBF (src/rNNGP.cpp:46-101).  theta = (sigma.sq, tau.sq, phi[, nu]).
    Returns (B, F, logDet)."""
    coords = _cm(coords)
    n = coords.size // 2
    B = np.zeros(max(n_indx(n, m), 1))
    F = np.zeros(n)
    th = np.zeros(4)
    th[:len(theta)] = theta
    ld = C.c_double()
    opts = _opts(precision)
    check(load().nngp_updateBF(B, F, coords, _i32(nnIndx), _i32(nnIndxLU), n,
                               m, th, cov_index(covModel), C.byref(ld),
                               C.byref(opts)))
    return B, F, ld.value


def updateConjBF(coords, nnIndx, nnIndxLU, m: int, phi, alpha, nu, covModel):
    """updateConjBF (src/cNNGP.cpp:24-69).  Returns (B, F)."""
    coords = _cm(coords)
    n = coords.size // 2
    B = np.zeros(max(n_indx(n, m), 1))
    F = np.zeros(n)
    opts = _opts()
    check(load().nngp_updateConjBF(B, F, coords, _i32(nnIndx), _i32(nnIndxLU),
                                   n, m, float(phi), float(alpha), float(nu),
                                   cov_index(covModel), C.byref(opts)))
    return B, F


def Q(B, F, u, v, nnIndx, nnIndxLU, m: int) -> float:
    """Q (src/util.cpp:261-280)."""
    F = _f64(F)
    q = C.c_double()
    check(load().nngp_Q(_f64(B), F, _f64(u), _f64(v), F.size, m, _i32(nnIndx),
                        _i32(nnIndxLU), C.byref(q)))
    return q.value


def prior_draw(B, F, nnIndx, nnIndxLU, z):
    """Sequential NNGP draw w_i = B_i.w_N(i) + sqrt(F_i) z_i
    (updateRep, src/rNNGP.cpp:24-42)."""
    F = _f64(F)
    w = np.zeros(F.size)
    check(load().nngp_prior_draw(_f64(B), F, _i32(nnIndx), _i32(nnIndxLU),
                                 F.size, _f64(z), w))
    return w


# ---------------------------------------------------------------------------
# drivers
# ---------------------------------------------------------------------------
class RRng:
    """R's default random number generator (Mersenne-Twister with
    inversion normals) seeded as set.seed(seed) seeds it, running in
    libnngp_hip.so.  Pass it as rng= to the samplers to follow R's stream;
    unif(k) / norm(k) draw like R's runif(k) / rnorm(k)."""

    def __init__(self, seed):
        self._s = _lib.NngpRng()
        check(load().nngp_rng_r_create(int(seed) & 0xFFFFFFFF,
                                       C.byref(self._s)))

    def struct(self, cls=None):
        """The {state, norm_rand, unif_rand} record, as the library's
        nngp_rng or as any structure `cls` with that layout."""
        if cls is None or cls is _lib.NngpRng:
            return self._s
        return cls(self._s.state, self._s.norm_rand, self._s.unif_rand)

    def unif(self, k=1):
        return np.array([self._s.unif_rand(self._s.state) for _ in range(k)])

    def norm(self, k=1):
        return np.array([self._s.norm_rand(self._s.state) for _ in range(k)])

    def close(self):
        if self._s.state:
            load().nngp_rng_r_destroy(C.byref(self._s))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _rng_struct(rng):
    if rng is None:
        return None
    return C.byref(rng.struct(NngpRng))


def rNNGP(y, X, p, n, m, coords, covModel, nnIndx, nnIndxLU, sigmaSqIG,
          tauSqIG, phiUnif, nuUnif, betaStarting, sigmaSqStarting,
          tauSqStarting, phiStarting, nuStarting, sigmaSqTuning, tauSqTuning,
          phiTuning, nuTuning, nSamples, nThreads=1, verbose=0, nReport=100,
          nRep=0, repIndx=None, rng=None, precision=0, seed=1,
          interrupt=None):
    """rNNGP (src/rNNGP.cpp:105-420), arguments in the .Call order of
    R/spNNGP.R:401-405.  rng: object with .struct(cls) giving the
    {state, norm_rand, unif_rand} callbacks, or None for the built-in
    generator seeded by `seed`.  interrupt: optional callable polled once
    per iteration (R_CheckUserInterrupt, src/rNNGP.cpp:377)."""
    cm = cov_index(covModel)
    nTheta = 4 if cm == 2 else 3
    beta = np.zeros(p * nSamples)
    theta = np.zeros(nTheta * nSamples)
    acc = C.c_int()
    rep = np.zeros(max(n * nRep, 1))
    ri = _i32(repIndx if repIndx is not None else np.zeros(max(nSamples, 1)))
    opts = _opts(precision, 0, seed, interrupt=interrupt)
    nu_u = _f64(nuUnif if nuUnif is not None else [0.0, 0.0])
    check(load().nngp_rNNGP(
        _f64(y), _cm(X), p, n, m, _cm(coords), cm, _i32(nnIndx),
        _i32(nnIndxLU), _f64(sigmaSqIG), _f64(tauSqIG), _f64(phiUnif),
        nu_u.ctypes.data_as(C.c_void_p), _f64(betaStarting),
        sigmaSqStarting, tauSqStarting, phiStarting, nuStarting,
        sigmaSqTuning, tauSqTuning, phiTuning, nuTuning, nSamples, nThreads,
        verbose, nReport, nRep, ri.ctypes.data_as(C.c_void_p),
        _rng_struct(rng), C.byref(opts), beta, theta,
        rep.ctypes.data_as(C.c_void_p), C.byref(acc)))
    out = {"p.beta.samples": beta.reshape(nSamples, p).T.copy(),
           "p.theta.samples": theta.reshape(nSamples, nTheta).T.copy(),
           "accept": acc.value}
    if nRep:
        out["y.rep.samples"] = rep.reshape(nRep, n).T.copy()
    return out


def mkUIndx(n, m, nnIndx, nnIndxLU, search_type=2):
    """mkUIndx (src/nn.cpp:25-56), R/nn.R:1-24: the sites holding each site
    as a neighbor (u.indx, u.indx.lu) and its slot in their lists (ui.indx).
    search_type 0 and 2 list them ascending, 1 descending."""
    ni = n_indx(n, m)
    u = np.zeros(max(ni, 1), np.int32)
    ui = np.zeros(max(ni, 1), np.int32)
    ulu = np.zeros(2 * n, np.int32)
    check(load().nngp_mkUIndx(n, m, _i32(nnIndx), u, ulu, ui, _i32(nnIndxLU),
                              int(search_type)))
    return {"u.indx": u[:ni], "u.indx.lu": ulu, "ui.indx": ui[:ni]}


SWEEPS = {"site": 0, "coloured": 1}


def sNNGP(y, X, p, n, m, coords, covModel, nnIndx, nnIndxLU, uIndx, uIndxLU,
          uiIndx, sigmaSqIG, tauSqIG, phiUnif, nuUnif, betaStarting,
          sigmaSqStarting, tauSqStarting, phiStarting, nuStarting,
          sigmaSqTuning, tauSqTuning, phiTuning, nuTuning, nSamples,
          nThreads=1, verbose=0, nReport=100, rng=None, seed=1,
          sweep="site"):
    """sNNGP (src/sNNGP.cpp:90-482), arguments in the .Call order of
    R/spNNGP.R:409-413.  sweep: "site" (the reference's order 0..n-1) or
    "coloured" (colour classes of the moral graph).  rng as for rNNGP; None:
    device Philox normals for the sweep and the built-in generator for the
    rest, seeded by `seed`."""
    cm = cov_index(covModel)
    nTheta = 4 if cm == 2 else 3
    beta = np.zeros(p * nSamples)
    theta = np.zeros(nTheta * nSamples)
    w = np.zeros(max(n * nSamples, 1))
    acc = C.c_int()
    opts = _opts(0, 0, seed, sweep=SWEEPS[sweep] if isinstance(sweep, str)
                 else int(sweep))
    nu_u = _f64(nuUnif if nuUnif is not None else [0.0, 0.0])
    check(load().nngp_sNNGP(
        _f64(y), _cm(X), p, n, m, _cm(coords), cm, _i32(nnIndx),
        _i32(nnIndxLU), _i32(uIndx), _i32(uIndxLU), _i32(uiIndx),
        _f64(sigmaSqIG), _f64(tauSqIG), _f64(phiUnif),
        nu_u.ctypes.data_as(C.c_void_p), _f64(betaStarting),
        sigmaSqStarting, tauSqStarting, phiStarting, nuStarting,
        sigmaSqTuning, tauSqTuning, phiTuning, nuTuning, nSamples, nThreads,
        verbose, nReport, _rng_struct(rng), C.byref(opts), beta, theta, w,
        C.byref(acc)))
    return {"p.beta.samples": beta.reshape(nSamples, p).T.copy(),
            "p.theta.samples": theta.reshape(nSamples, nTheta).T.copy(),
            "p.w.samples": w[:n * nSamples].reshape(nSamples, n).T.copy(),
            "accept": acc.value}


def cNNGP(y, X, p, n, coords, thetaAlpha, nnIndx, nnIndxLU, X0, coords0, n0,
          nnIndx0, g, m, sigmaSqIG, covModel, nThreads=1, verbose=0,
          precision=0, interrupt=None):
    """cNNGP (src/cNNGP.cpp:73-358), .Call order of R/spConjNNGP.R:336.
    thetaAlpha: nTheta x g (rows phi, alpha[, nu]); nnIndx0: n0 x m.
    interrupt: optional callable polled once per grid point."""
    cm = cov_index(covModel)
    betaHat = np.zeros(p * g)
    ab = np.zeros(2 * g)
    bBInv = np.zeros(p * p * g)
    bb = np.zeros(p * g)
    y0 = np.zeros(max(n0 * g, 1))
    y0v = np.zeros(max(n0 * g, 1))
    opts = _opts(precision, interrupt=interrupt)
    keep = []

    def vp(a, conv):
        if a is None or n0 == 0:
            return None
        arr = conv(a)
        keep.append(arr)
        return arr.ctypes.data_as(C.c_void_p)

    check(load().nngp_cNNGP(
        _f64(y), _cm(X), p, n, _cm(coords), _cm(thetaAlpha), _i32(nnIndx),
        _i32(nnIndxLU), vp(X0, _cm), vp(coords0, _cm), n0,
        vp(nnIndx0, lambda a: _i32(np.asarray(a).T.reshape(-1))
           if np.asarray(a).ndim == 2 else _i32(a)),
        g, m, _f64(sigmaSqIG), cm, nThreads, verbose, C.byref(opts), betaHat,
        ab, bBInv, bb, y0.ctypes.data_as(C.c_void_p),
        y0v.ctypes.data_as(C.c_void_p)))
    return {"beta.hat": betaHat.reshape(g, p).T.copy(),
            "ab": ab.reshape(g, 2).T.copy(),
            "bB.inv": bBInv.reshape(g, p * p).T.copy(),
            "bb": bb.reshape(g, p).T.copy(),
            "y.0.hat": y0[:n0 * g].reshape(g, n0).T.copy(),
            "y.0.hat.var": y0v[:n0 * g].reshape(g, n0).T.copy()}


def rNNGPPredict(X, y, coords, n, p, m, X0, coords0, q, nnIndx0,
                 betaSamples, thetaSamples, nSamples, covModel, nThreads=1,
                 verbose=0, nReport=100, z=None, rng=None, seed=1,
                 precision=0):
    """rNNGPPredict (src/rNNGPPredict.cpp:25-222), .Call order of
    R/generics.R:674.  betaSamples p x S, thetaSamples nTheta x S, nnIndx0
    q x m (0-based).  z: q*S normals with z[i*S + s], or None."""
    out = np.zeros(q * nSamples)
    opts = _opts(precision, 0, seed)
    nn0 = np.asarray(nnIndx0)
    nn0 = _i32(nn0.T.reshape(-1)) if nn0.ndim == 2 else _i32(nn0)
    zz = None if z is None else _f64(z)
    check(load().nngp_rNNGPPredict(
        _cm(X), _f64(y), _cm(coords), n, p, m, _cm(X0), _cm(coords0), q, nn0,
        _cm(betaSamples), _cm(thetaSamples), nSamples, cov_index(covModel),
        nThreads, verbose, nReport,
        None if zz is None else zz.ctypes.data_as(C.c_void_p),
        _rng_struct(rng), C.byref(opts), out))
    return {"p.y.0": out.reshape(nSamples, q).T.copy()}


def sNNGPPredict(X, y, coords, n, p, m, X0, coords0, q, nnIndx0,
                 betaSamples, thetaSamples, wSamples, nSamples, covModel,
                 family=1, nThreads=1, verbose=0, nReport=100, z=None,
                 rng=None, seed=1):
    """sNNGPPredict (src/sNNGPPredict.cpp:25-241), .Call order of
    R/generics.R:671.  wSamples n x S (the latent fit's p.w.samples);
    z: 2*q*S normals (w0's then y0's, each z[i*S + s]) or None."""
    out = np.zeros(q * nSamples)
    w0 = np.zeros(q * nSamples)
    opts = _opts(0, 0, seed)
    nn0 = np.asarray(nnIndx0)
    nn0 = _i32(nn0.T.reshape(-1)) if nn0.ndim == 2 else _i32(nn0)
    zz = None if z is None else _f64(z)
    check(load().nngp_sNNGPPredict(
        _cm(X), None, _cm(coords), n, p, m, _cm(X0), _cm(coords0), q, nn0,
        _cm(betaSamples), _cm(thetaSamples), _cm(wSamples), nSamples,
        int(family), cov_index(covModel), nThreads, verbose, nReport,
        None if zz is None else zz.ctypes.data_as(C.c_void_p),
        _rng_struct(rng), C.byref(opts), out, w0))
    return {"p.y.0": out.reshape(nSamples, q).T.copy(),
            "p.w.0": w0.reshape(nSamples, q).T.copy()}


def rNNGPReplicated(X, p, n, m, coords, covModel, nnIndx, nnIndxLU, beta,
                    theta, nSamples, nThreads=1, verbose=0, nReport=100,
                    rng=None, seed=1):
    """rNNGPReplicated (src/rNNGPReplicated.cpp:78-230), .Call order of
    R/generics.R:147.  X n x p, coords n x 2 in NNGP order, beta nSamples x
    p, theta nTheta x nSamples.  rng: as for rNNGP (normals in the
    reference's order); None: Philox normals keyed by (seed, site,
    sample).  Returns {'y.rep.samples': n x nSamples}."""
    del nThreads  # accepted for signature parity; the device runs it
    out = np.zeros(max(n * nSamples, 1))
    opts = _opts(0, 0, seed)
    Xc = _cm(X) if p else None
    bc = _cm(np.asarray(beta, np.float64).reshape(nSamples, p)) if p else None
    check(load().nngp_rNNGPReplicated(
        None if Xc is None else Xc.ctypes.data_as(C.c_void_p), p, n, m,
        _cm(coords), cov_index(covModel), _i32(nnIndx), _i32(nnIndxLU),
        None if bc is None else bc.ctypes.data_as(C.c_void_p), _cm(theta),
        nSamples, verbose, nReport, _rng_struct(rng), C.byref(opts), out))
    # n x nSamples as a column-major view of the library's output (no copy
    # of what can be gigabytes)
    return {"y.rep.samples": out[:n * nSamples].reshape(nSamples, n).T}


# ---------------------------------------------------------------------------
# device-resident model
# ---------------------------------------------------------------------------
class Model:
    """Coordinates, neighbor index, X and y kept in HBM (nngp_model_*)."""

    def __init__(self, coords, nnIndx, nnIndxLU, m, covModel, X=None, y=None,
                 precision=0, shape=None):
        coords = _cm(coords)
        self.n = coords.size // 2
        self.m = m
        self.cov = cov_index(covModel)
        self.p = 0 if X is None else np.asarray(X).shape[1]
        self._h = C.c_void_p()
        self._opts = _opts(precision, shape=shape)
        Xc = None if X is None else _cm(X)
        yc = None if y is None else _f64(y)
        self._keep = (coords, Xc, yc)
        check(load().nngp_model_create(
            C.byref(self._h), self.n, m, self.p, coords, _i32(nnIndx),
            _i32(nnIndxLU),
            None if Xc is None else Xc.ctypes.data_as(C.c_void_p),
            None if yc is None else yc.ctypes.data_as(C.c_void_p), self.cov,
            C.byref(self._opts)))

    def close(self):
        if self._h:
            load().nngp_model_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, theta, slot=0, beta=None, want_xprod=True):
        th = np.zeros(4)
        th[:len(theta)] = theta
        out = NngpPassOut()
        b = None if beta is None else _f64(beta)
        check(load().nngp_model_update(
            self._h, th, slot,
            None if b is None else b.ctypes.data_as(C.c_void_p),
            int(want_xprod), C.byref(out)))
        p = self.p
        res = {"logDet": out.logDet, "qResid": out.qResid}
        if want_xprod and p:
            X = np.array(out.XtViX[:p * p]).reshape(p, p).T
            res["XtViX"] = np.tril(X) + np.tril(X, -1).T
            res["XtViy"] = np.array(out.XtViy[:p])
            res["ytViy"] = out.ytViy
        return res

    def profile(self, enable=True, every=1):
        """Start (and zero) or stop per-kernel event timing and the loop
        clock.  every=k > 1 times every k-th updateBF launch (and the sum
        launch after it) instead of each one."""
        check(load().nngp_model_profile(self._h,
                                        max(int(every), 1) if enable else 0))

    def timing(self):
        """{'bf_ms', 'bf_launches', 'sum_ms', 'sum_launches', 'loop_ms',
        'loop_iters'} so far: kernel times by HIP events, and the host wall
        time of the samplers' iteration loops (first proposal to last
        decision, stream drained at both ends; the initial pass outside)."""
        t = np.zeros(4)
        check(load().nngp_model_timing(self._h, t))
        lt = np.zeros(2)
        check(load().nngp_model_loop_timing(self._h, lt))
        return {"bf_ms": t[0], "bf_launches": int(t[1]), "sum_ms": t[2],
                "sum_launches": int(t[3]), "loop_ms": lt[0],
                "loop_iters": int(lt[1])}

    def autotune(self, theta):
        """Time the updateBF launch shapes at theta and keep the fastest.
        Returns [(kind, threads, ms), ...] and the chosen (kind, threads)."""
        th = np.zeros(4)
        th[:len(theta)] = theta
        r = _lib.NngpTuneReport()
        check(load().nngp_model_autotune(self._h, th, C.byref(r)))
        return ([(_shape_name(r.family[t], r.mode[t]), r.threads[t], r.ms[t])
                 for t in range(r.n)],
                (_shape_name(r.best_family, r.best_mode), r.best_threads))

    def set_shape(self, family, threads, mode=0):
        """Pin the updateBF launch shape: family 0 (one site per lane), 1
        (per wavefront) or 2 (per 16-lane row: four sites a wavefront, m <=
        32, FP64), threads per block, and mode as in the autotuner's
        report (1 and 3: the mixed per-lane kernel's covariance copy in LDS
        or in registers; 2: the FP64 per-lane kernel's covariate rows staged
        in LDS)."""
        check(load().nngp_model_set_shape(self._h, int(family), int(threads),
                                          int(mode)))

    def qresid(self, beta, slot=0):
        q = C.c_double()
        check(load().nngp_model_qresid(self._h, slot, _f64(beta), C.byref(q)))
        return q.value

    def get_bf(self, slot=0):
        B = np.zeros(max(n_indx(self.n, self.m), 1))
        F = np.zeros(self.n)
        check(load().nngp_model_get_bf(self._h, slot,
                                       B.ctypes.data_as(C.c_void_p),
                                       F.ctypes.data_as(C.c_void_p)))
        return B, F

    def rNNGP(self, sigmaSqIG, tauSqIG, phiUnif, nuUnif, betaStarting,
              sigmaSqStarting, tauSqStarting, phiStarting, nuStarting,
              sigmaSqTuning, tauSqTuning, phiTuning, nuTuning, nSamples,
              verbose=0, nReport=100, rng=None, seed=1):
        p = self.p
        nTheta = 4 if self.cov == 2 else 3
        beta = np.zeros(p * nSamples)
        theta = np.zeros(nTheta * nSamples)
        acc = C.c_int()
        opts = _opts(0, 0, seed)
        nu_u = _f64(nuUnif if nuUnif is not None else [0.0, 0.0])
        check(load().nngp_model_rNNGP(
            self._h, _f64(sigmaSqIG), _f64(tauSqIG), _f64(phiUnif),
            nu_u.ctypes.data_as(C.c_void_p), _f64(betaStarting),
            sigmaSqStarting, tauSqStarting, phiStarting, nuStarting,
            sigmaSqTuning, tauSqTuning, phiTuning, nuTuning, nSamples,
            verbose, nReport, _rng_struct(rng), C.byref(opts), beta, theta,
            C.byref(acc)))
        return {"p.beta.samples": beta.reshape(nSamples, p).T.copy(),
                "p.theta.samples": theta.reshape(nSamples, nTheta).T.copy(),
                "accept": acc.value}


# ---------------------------------------------------------------------------
# kriging server
# ---------------------------------------------------------------------------
class Server:
    """Resident kriging service (nngp_server_*): the fitted rows, their
    neighbor-search index and the posterior samples stay in HBM; requests
    are batched continuously as they arrive.  Thread-safe: submit and wait
    may be called from different threads."""

    def __init__(self, X, y, coords, m, betaSamples, thetaSamples, covModel,
                 device_bytes=0, host_bytes=0, batch_sites=0, depth=0,
                 device=0):
        X = np.asarray(X, dtype=np.float64)
        self.n, self.p = X.shape
        self.m = m
        self.cov = cov_index(covModel)
        bs = np.asarray(betaSamples, dtype=np.float64)
        self.S = bs.shape[1]
        o = _lib.NngpServerOpts(device_bytes, host_bytes, batch_sites, depth,
                                device, (C.c_int * 8)())
        self._h = C.c_void_p()
        check(load().nngp_server_create(
            C.byref(self._h), _cm(X), _f64(y), _cm(coords), self.n, self.p,
            m, _cm(bs), _cm(thetaSamples), self.S, self.cov, C.byref(o)))

    def close(self):
        if self._h:
            load().nngp_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, X0, coords0, seed) -> int:
        X0 = np.asarray(X0, dtype=np.float64)
        t = C.c_longlong()
        check(load().nngp_server_submit(self._h, X0.shape[0], _cm(X0),
                                        _cm(coords0), int(seed), C.byref(t)))
        return t.value

    def wait(self, ticket=-1, timeout=-1.0, copy=True):
        """Wait for `ticket` (or any finished request when -1).  Returns
        (ticket, info dict); info['y0'] is the q x S result (a copy when
        copy=True, else None) and the ticket still has to be released."""
        t = C.c_longlong(ticket)
        info = _lib.NngpRequestInfo()
        check(load().nngp_server_wait(self._h, C.byref(t), float(timeout),
                                      C.byref(info)))
        res = {"q": info.q, "status": info.status, "t_submit": info.t_submit,
               "t_admit": info.t_admit, "t_done": info.t_done, "y0": None}
        if copy and info.status == 0:
            a = np.ctypeslib.as_array(info.y0, shape=(info.nSamples, info.q))
            res["y0"] = a.T.copy()
        return t.value, res

    def release(self, ticket):
        check(load().nngp_server_release(self._h, int(ticket)))

    def predict(self, X0, coords0, seed):
        """Synchronous request: {'p.y.0': q x S}."""
        X0 = np.asarray(X0, dtype=np.float64)
        q = X0.shape[0]
        out = np.zeros(q * self.S)
        check(load().nngp_server_predict(self._h, q, _cm(X0), _cm(coords0),
                                         int(seed), out))
        return {"p.y.0": out.reshape(self.S, q).T.copy()}

    def stats(self):
        s = _lib.NngpServerStats()
        check(load().nngp_server_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}


# ---------------------------------------------------------------------------
# conjugate k-fold grid search
# ---------------------------------------------------------------------------
class CvGrid:
    """spConjNNGP's k-fold scoring of a theta grid (nngp_cvgrid_*): folds
    fixed at construction (ho: k x nHo held-out rows, 0-based), every run
    scores all folds x grid points in one graph launch."""

    def __init__(self, y, X, coords, m, ho, sigmaSqIG, covModel, shape=None,
                 search_type="cb"):
        """search_type: each fold's neighbor search, "cb" (mkNNIndxCB) or
        "brute" (mkNNIndx), R's search.type."""
        if search_type not in ("cb", "brute"):
            raise ValueError(f"search_type must be 'cb' or 'brute', "
                             f"not {search_type!r}")
        X = np.asarray(X, dtype=np.float64)
        self.n, self.p = X.shape
        ho = np.asarray(ho)
        self.k, self.nho = ho.shape
        self.m = m
        self.cov = cov_index(covModel)
        self._h = C.c_void_p()
        opts = _opts(shape=shape, search=int(search_type == "brute"))
        check(load().nngp_cvgrid_create(
            C.byref(self._h), _f64(y), _cm(X), self.p, self.n, _cm(coords),
            m, self.k, _i32(ho.reshape(-1)), self.nho, _f64(sigmaSqIG),
            self.cov, C.byref(opts)))

    def close(self):
        if self._h:
            load().nngp_cvgrid_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, thetaAlpha, details=False, predictions=False):
        """thetaAlpha: g x nTheta (columns phi, alpha[, nu]).  Returns
        {'scores': g x 2 (rmspe, crps), 'graph_ms', 'bf_config', 'tune'}
        and, on request, per (fold, point) beta.hat, ab, bB.inv and the
        held-out predictions (arrays indexed [fold, point, ...])."""
        ta = np.asarray(thetaAlpha, dtype=np.float64)
        g = ta.shape[0]
        k, p, nho = self.k, self.p, self.nho
        scores = np.zeros(2 * g)
        d = _lib.NngpCvGridDetail()
        keep = []

        def buf(size):
            a = np.zeros(size)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        if details:
            d.betaHat = buf(p * g * k)
            d.ab = buf(2 * g * k)
            d.bBInv = buf(p * p * g * k)
        if predictions:
            d.y0Hat = buf(nho * g * k)
            d.y0HatVar = buf(nho * g * k)
        check(load().nngp_cvgrid_run(self._h, _f64(ta.reshape(-1)), g, scores,
                                     C.byref(d)))
        out = {"scores": scores.reshape(2, g).T.copy(),
               "graph_ms": d.graph_ms,
               "bf_config": (_shape_name(d.bf_family, d.bf_mode),
                             d.bf_threads),
               "tune": [(_shape_name(d.tune_family[t], d.tune_mode[t]),
                         d.tune_threads[t], d.tune_ms[t])
                        for t in range(d.tune_n)]}
        i = 0
        if details:
            out["beta.hat"] = keep[0].reshape(k, g, p)
            out["ab"] = keep[1].reshape(k, g, 2)
            out["bB.inv"] = keep[2].reshape(k, g, p, p).transpose(0, 1, 3, 2)
            i = 3
        if predictions:
            out["y.0.hat"] = keep[i].reshape(k, g, nho)
            out["y.0.hat.var"] = keep[i + 1].reshape(k, g, nho)
        return out
