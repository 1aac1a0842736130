"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

Function names follow the reference routines they restate (see
nngp_oracle.h for the file:line of each).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NNGP_ORACLE_LIB names the sanitized build (`make -C oracle sanitize`)
_LIB_PATH = os.environ.get("NNGP_ORACLE_LIB") or os.path.join(
    _HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class OrRng(C.Structure):
    _fields_ = [
        ("state", C.c_void_p),
        ("norm_rand", C.CFUNCTYPE(C.c_double, C.c_void_p)),
        ("unif_rand", C.CFUNCTYPE(C.c_double, C.c_void_p)),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_last_error.restype = C.c_char_p
        L.or_getNNIndx.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
        L.or_dist2.restype = C.c_double
        L.or_dist2.argtypes = [C.c_double] * 4
        L.or_bessel_k.restype = C.c_double
        L.or_bessel_k.argtypes = [C.c_double, C.c_double]
        L.or_spCor.restype = C.c_double
        L.or_spCor.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.or_rsort_with_index.argtypes = [_dp, _ip, C.c_int]
        for f in (L.or_mkNNIndx, L.or_mkNNIndxCB):
            f.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp, _ip, C.c_int]
        L.or_nn2.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, _ip, _dp]
        L.or_potrf_L.argtypes = [C.c_int, _dp]
        L.or_potri_L.argtypes = [C.c_int, _dp]
        L.or_symv_L.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_updateBF.argtypes = [_dp, _dp, _dp, _ip, _ip, C.c_int, C.c_int,
                                  _dp, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.or_updateConjBF.argtypes = [_dp, _dp, _dp, _ip, _ip, C.c_int,
                                      C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int]
        L.or_Q.restype = C.c_double
        L.or_Q.argtypes = [_dp, _dp, _dp, _dp, C.c_int, _ip, _ip]
        L.or_rNNGP.argtypes = (
            [_dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip, _ip,
             _dp, _dp, _dp, _dp, _dp]
            + [C.c_double] * 8
            + [C.c_int, C.c_int, C.c_int, _vp, C.POINTER(OrRng), _dp, _dp,
               _vp, C.POINTER(C.c_int)])
        L.or_rNNGPPredict.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int,
                                      C.c_int, _dp, _dp, C.c_int, _ip, _dp,
                                      _dp, C.c_int, C.c_int, C.c_int, _dp,
                                      _dp]
        L.or_cNNGP.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp, _dp, _ip, _ip,
                               _dp, _dp, C.c_int, _ip, C.c_int, C.c_int, _dp,
                               C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_nngp_draw.argtypes = [_dp, _dp, _ip, _ip, C.c_int, _dp, _dp]
        L.or_rNNGPReplicated.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp,
                                         C.c_int, _ip, _ip, _dp, _dp, C.c_int,
                                         C.c_int, C.POINTER(OrRng), _dp]
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint), C.c_uint,
                                       C.c_uint]
        L.or_philox_normal.restype = C.c_double
        L.or_philox_normal.argtypes = [C.c_ulonglong, C.c_uint, C.c_uint]
        L.or_sNNGPPredict.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int,
                                      _dp, C.c_int, _ip, _dp, _dp, _dp,
                                      C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_mkUIndx.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip,
                                 C.c_int]
        L.or_rgamma.restype = C.c_double
        L.or_rgamma.argtypes = [C.POINTER(OrRng), C.c_double, C.c_double]
        L.or_sNNGP.argtypes = (
            [_dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip, _ip,
             _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
            + [C.c_double] * 6
            + [C.c_int, C.c_int, _vp, C.POINTER(OrRng), _dp, _dp, _dp,
               C.POINTER(C.c_int)])
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise RuntimeError(lib().or_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def n_indx(n: int, m: int) -> int:
    return int((1 + m) / 2 * m + (n - m - 1) * m)


def getNNIndx(i: int, m: int):
    a, b = C.c_int(), C.c_int()
    lib().or_getNNIndx(i, m, C.byref(a), C.byref(b))
    return a.value, b.value


def bessel_k(x: float, nu: float) -> float:
    return lib().or_bessel_k(float(x), float(nu))


def spCor(D: float, phi: float, nu: float, covModel: int) -> float:
    return lib().or_spCor(float(D), float(phi), float(nu), int(covModel))


def rsort_with_index(x, idx):
    x = _f64(x).copy()
    idx = _i32(idx).copy()
    lib().or_rsort_with_index(x, idx, len(x))
    return x, idx


def _nn(fn, coords, m, nThreads):
    coords = _f64(np.asarray(coords, dtype=np.float64).T.reshape(-1)) \
        if np.asarray(coords).ndim == 2 else _f64(coords)
    n = coords.size // 2
    ni = n_indx(n, m)
    nnIndx = np.zeros(ni, np.int32)
    nnDist = np.zeros(ni, np.float64)
    nnIndxLU = np.zeros(2 * n, np.int32)
    _check(fn(n, m, coords, nnIndx, nnDist, nnIndxLU, nThreads))
    return {"nnIndx": nnIndx, "nnDist": nnDist, "nnIndxLU": nnIndxLU}


def mkNNIndx(coords, m, nThreads=1):
    """coords: n x 2 array (or flat column-major 2n vector)."""
    return _nn(lib().or_mkNNIndx, coords, m, nThreads)


def mkNNIndxCB(coords, m, nThreads=1):
    return _nn(lib().or_mkNNIndxCB, coords, m, nThreads)


def nn2(coords, coords0, k):
    coords = np.asarray(coords, np.float64)
    coords0 = np.asarray(coords0, np.float64)
    n, q = coords.shape[0], coords0.shape[0]
    idx = np.zeros(q * k, np.int32)
    dist = np.zeros(q * k, np.float64)
    _check(lib().or_nn2(n, _f64(coords.T.reshape(-1)), q,
                        _f64(coords0.T.reshape(-1)), k, idx, dist))
    return idx.reshape(k, q).T.copy(), dist.reshape(k, q).T.copy()


def potrf_L(A):
    A = np.asfortranarray(A, dtype=np.float64).copy()
    flat = _f64(A.T.reshape(-1))
    info = lib().or_potrf_L(A.shape[0], flat)
    return flat.reshape(A.shape[0], A.shape[0]).T.copy(), info


def potri_L(L):
    n = L.shape[0]
    flat = _f64(np.asarray(L, np.float64).T.reshape(-1))
    info = lib().or_potri_L(n, flat)
    return flat.reshape(n, n).T.copy(), info


def symv_L(A, x):
    n = A.shape[0]
    y = np.zeros(n)
    lib().or_symv_L(n, _f64(np.asarray(A, np.float64).T.reshape(-1)),
                    _f64(x), y)
    return y


def updateBF(coords_flat, nnIndx, nnIndxLU, n, m, theta, covModel,
             nThreads=1):
    B = np.zeros(max(n_indx(n, m), 1))
    F = np.zeros(n)
    ld = C.c_double()
    th = np.zeros(4)
    th[:len(theta)] = theta
    _check(lib().or_updateBF(B, F, _f64(coords_flat), _i32(nnIndx),
                             _i32(nnIndxLU), n, m, th, covModel, nThreads,
                             C.byref(ld)))
    return B, F, ld.value


def updateConjBF(coords_flat, nnIndx, nnIndxLU, n, m, phi, alpha, nu,
                 covModel, nThreads=1):
    B = np.zeros(max(n_indx(n, m), 1))
    F = np.zeros(n)
    _check(lib().or_updateConjBF(B, F, _f64(coords_flat), _i32(nnIndx),
                                 _i32(nnIndxLU), n, m, phi, alpha, nu,
                                 covModel, nThreads))
    return B, F


def Q(B, F, u, v, nnIndx, nnIndxLU):
    n = len(F)
    return lib().or_Q(_f64(B), _f64(F), _f64(u), _f64(v), n, _i32(nnIndx),
                      _i32(nnIndxLU))


def nngp_draw(B, F, nnIndx, nnIndxLU, z):
    n = len(F)
    w = np.zeros(n)
    lib().or_nngp_draw(_f64(B), _f64(F), _i32(nnIndx), _i32(nnIndxLU), n,
                       _f64(z), w)
    return w


def rNNGPReplicated(X, p, n, m, coords_flat, covModel, nnIndx, nnIndxLU,
                    beta, theta, nSamples, tape, nThreads=1):
    """src/rNNGPReplicated.cpp:78-230.  X n x p and beta nSamples x p
    column-major flat, theta nTheta x nSamples flat; normals from the
    TapeRng `tape` in the reference's order.  Returns y.rep, n x nSamples."""
    y = np.zeros(n * nSamples)
    r = tape.struct(OrRng)
    _check(lib().or_rNNGPReplicated(_f64(X), p, n, m, _f64(coords_flat),
                                    covModel, _i32(nnIndx), _i32(nnIndxLU),
                                    _f64(beta), _f64(theta), nSamples,
                                    nThreads, C.byref(r), y))
    return y.reshape(nSamples, n).T.copy()


class TapeRng:
    """Serves pre-drawn standard normals and uniforms in call order.

    Two TapeRng objects built from the same arrays give two identical
    streams, which is how the oracle and the product are fed the same
    random numbers.  Exhausting a tape is an error (returns NaN)."""

    def __init__(self, normals, uniforms):
        self.normals = np.ascontiguousarray(normals, np.float64)
        self.uniforms = np.ascontiguousarray(uniforms, np.float64)
        self.i_n = 0
        self.i_u = 0
        NF = C.CFUNCTYPE(C.c_double, C.c_void_p)

        def _norm(_):
            if self.i_n >= self.normals.size:
                return float("nan")
            v = self.normals[self.i_n]
            self.i_n += 1
            return float(v)

        def _unif(_):
            if self.i_u >= self.uniforms.size:
                return float("nan")
            v = self.uniforms[self.i_u]
            self.i_u += 1
            return float(v)

        self._cb_n = NF(_norm)
        self._cb_u = NF(_unif)

    def struct(self, cls=OrRng):
        # CFUNCTYPE types are cached by signature, so the callbacks fit any
        # structure with the same {state, norm_rand, unif_rand} layout.
        return cls(None, self._cb_n, self._cb_u)


def rNNGP(y, X, p, n, m, coords_flat, covModel, nnIndx, nnIndxLU, sigmaSqIG,
          tauSqIG, phiUnif, nuUnif, betaStarting, sigmaSqStarting,
          tauSqStarting, phiStarting, nuStarting, sigmaSqTuning, tauSqTuning,
          phiTuning, nuTuning, nSamples, rng: TapeRng, nThreads=1, nRep=0,
          repIndx=None):
    nTheta = 4 if covModel == 2 else 3
    beta = np.zeros(p * nSamples)
    theta = np.zeros(nTheta * nSamples)
    acc = C.c_int()
    rs = rng.struct(OrRng)
    rep = np.zeros(max(n * nRep, 1))
    ri = _i32(repIndx if repIndx is not None else np.zeros(nSamples))
    st = lib().or_rNNGP(
        _f64(y), _f64(np.asarray(X, np.float64).T.reshape(-1)
                      if np.asarray(X).ndim == 2 else X),
        p, n, m, _f64(coords_flat), covModel, _i32(nnIndx), _i32(nnIndxLU),
        _f64(sigmaSqIG), _f64(tauSqIG), _f64(phiUnif),
        _f64(nuUnif if nuUnif is not None else [0.0, 0.0]),
        _f64(betaStarting), sigmaSqStarting, tauSqStarting, phiStarting,
        nuStarting, sigmaSqTuning, tauSqTuning, phiTuning, nuTuning,
        nSamples, nThreads, nRep,
        ri.ctypes.data_as(C.c_void_p), C.byref(rs), beta, theta,
        rep.ctypes.data_as(C.c_void_p), C.byref(acc))
    _check(st)
    out = {"p.beta.samples": beta.reshape(nSamples, p).T.copy(),
           "p.theta.samples": theta.reshape(nSamples, nTheta).T.copy(),
           "accept": acc.value}
    if nRep:
        out["y.rep.samples"] = rep.reshape(nRep, n).T.copy()
    return out


def rNNGPPredict(X, y, coords_flat, n, p, m, X0, coords0_flat, q, nnIndx0,
                 betaSamples, thetaSamples, nSamples, covModel, z,
                 nThreads=1):
    out = np.zeros(q * nSamples)
    _check(lib().or_rNNGPPredict(
        _f64(X), _f64(y), _f64(coords_flat), n, p, m, _f64(X0),
        _f64(coords0_flat), q, _i32(nnIndx0), _f64(betaSamples),
        _f64(thetaSamples), nSamples, covModel, nThreads, _f64(z), out))
    return out.reshape(nSamples, q).T.copy()


def cNNGP(y, X, p, n, coords_flat, thetaAlpha, nnIndx, nnIndxLU, X0,
          coords0_flat, n0, nnIndx0, g, m, sigmaSqIG, covModel, nThreads=1):
    betaHat = np.zeros(p * g)
    ab = np.zeros(2 * g)
    bBInv = np.zeros(p * p * g)
    bb = np.zeros(p * g)
    y0 = np.zeros(max(n0 * g, 1))
    y0v = np.zeros(max(n0 * g, 1))
    _check(lib().or_cNNGP(
        _f64(y), _f64(X), p, n, _f64(coords_flat), _f64(thetaAlpha),
        _i32(nnIndx), _i32(nnIndxLU), _f64(X0 if n0 else np.zeros(1)),
        _f64(coords0_flat if n0 else np.zeros(2)), n0,
        _i32(nnIndx0 if n0 else np.zeros(1)), g, m, _f64(sigmaSqIG),
        covModel, nThreads, betaHat, ab, bBInv, bb, y0, y0v))
    return {"beta.hat": betaHat.reshape(g, p).T.copy(),
            "ab": ab.reshape(g, 2).T.copy(),
            "bB.inv": bBInv.reshape(g, p * p).T.copy(),
            "bb": bb.reshape(g, p).T.copy(),
            "y.0.hat": y0[:n0 * g].reshape(g, n0).T.copy(),
            "y.0.hat.var": y0v[:n0 * g].reshape(g, n0).T.copy()}


def philox4x32_10(ctr, key):
    """Philox4x32-10 block of counter ctr[4] under key[2] (four uint32)."""
    c = (C.c_uint * 4)(*ctr)
    lib().or_philox4x32_10(c, key[0], key[1])
    return [int(v) for v in c]


def philox_normals(seed, q, S):
    """The kriging server's z for a request: q*S normals, z[i*S + s]."""
    L = lib()
    z = np.empty(q * S)
    for i in range(q):
        for s in range(S):
            z[i * S + s] = L.or_philox_normal(seed, i, s)
    return z


def mkUIndx(n, m, nnIndx, nnIndxLU, searchType=2):
    """src/nn.cpp:25-56: u.indx, u.indx.lu, ui.indx of R/nn.R:1-24."""
    ni = n_indx(n, m)
    u = np.zeros(max(ni, 1), np.int32)
    ui = np.zeros(max(ni, 1), np.int32)
    ulu = np.zeros(2 * n, np.int32)
    _check(lib().or_mkUIndx(n, m, _i32(nnIndx), u, ulu, ui, _i32(nnIndxLU),
                            searchType))
    return {"u.indx": u[:ni], "u.indx.lu": ulu, "ui.indx": ui[:ni]}


def rgamma(tape, a, scale, k=1):
    """k Gamma(a, scale) deviates by algorithm GD from the TapeRng tape."""
    r = tape.struct(OrRng)
    return np.array([lib().or_rgamma(C.byref(r), a, scale) for _ in range(k)])


def sNNGP(y, X, p, n, m, coords_flat, covModel, nnIndx, nnIndxLU, uIndx,
          uIndxLU, uiIndx, sigmaSqIG, tauSqIG, phiUnif, nuUnif, betaStarting,
          sigmaSqStarting, tauSqStarting, phiStarting, nuStarting, phiTuning,
          nuTuning, nSamples, rng: TapeRng, nThreads=1, visit=None):
    """src/sNNGP.cpp:90-482; X n x p (2-D) or column-major flat.  visit:
    the site order of the w sweep (None: 0..n-1, the reference's)."""
    nTheta = 4 if covModel == 2 else 3
    beta = np.zeros(p * nSamples)
    theta = np.zeros(nTheta * nSamples)
    w = np.zeros(max(n * nSamples, 1))
    acc = C.c_int()
    rs = rng.struct(OrRng)
    vis = None if visit is None else _i32(visit)
    Xf = _f64(np.asarray(X, np.float64).T.reshape(-1)
              if np.asarray(X).ndim == 2 else X)
    _check(lib().or_sNNGP(
        _f64(y), Xf, p, n, m, _f64(coords_flat), covModel, _i32(nnIndx),
        _i32(nnIndxLU), _i32(uIndx), _i32(uIndxLU), _i32(uiIndx),
        _f64(sigmaSqIG), _f64(tauSqIG), _f64(phiUnif),
        _f64(nuUnif if nuUnif is not None else [0.0, 0.0]),
        _f64(betaStarting), sigmaSqStarting, tauSqStarting, phiStarting,
        nuStarting, phiTuning, nuTuning, nSamples, nThreads,
        None if vis is None else vis.ctypes.data_as(C.c_void_p),
        C.byref(rs), beta, theta, w, C.byref(acc)))
    return {"p.beta.samples": beta.reshape(nSamples, p).T.copy(),
            "p.theta.samples": theta.reshape(nSamples, nTheta).T.copy(),
            "p.w.samples": w[:n * nSamples].reshape(nSamples, n).T.copy(),
            "accept": acc.value}


def sNNGPPredict(X0_flat, coords_flat, n, p, m, coords0_flat, q, nnIndx0,
                 beta, theta, w, nSamples, covModel, z):
    """src/sNNGPPredict.cpp:25-241.  Flat column-major inputs; nnIndx0
    q x m column-major flat; beta p x S, theta nTheta x S, w n x S flat;
    z 2 q S normals.  Returns (p.y.0, p.w.0), each q x S."""
    y0 = np.zeros(q * nSamples)
    w0 = np.zeros(q * nSamples)
    _check(lib().or_sNNGPPredict(_f64(X0_flat), _f64(coords_flat), n, p, m,
                                 _f64(coords0_flat), q, _i32(nnIndx0),
                                 _f64(beta), _f64(theta), _f64(w), nSamples,
                                 covModel, _f64(z), y0, w0))
    return (y0.reshape(nSamples, q).T.copy(), w0.reshape(nSamples, q).T.copy())
