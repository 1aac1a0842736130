"""spnngp_amd — MI355X (gfx950) NNGP hot path of spNNGP 1.0.1.

The compute lives in libnngp_hip.so (hand-written HIP kernels behind the C
ABI of include/nngp.h); this package is the numpy-facing host mirror of the
reference's native routines:

    mkNNIndxCB, mkNNIndx, nn2          neighbor search
    mkUIndx                            sites holding each site as a neighbor
    updateBF, updateConjBF, Q          per-site systems and quadratic forms
    rNNGP, sNNGP, cNNGP, rNNGPPredict,
    sNNGPPredict                       the .Call drivers
    Model                              device-resident fitted data
    Server                             resident kriging service
    CvGrid                             conjugate k-fold grid search
    spNNGP, spConjNNGP, predict,       the R-level front ends
    crps, rmspe

There is no CPU fallback: without the built library every call raises.
"""
from ._lib import (NNGP_ERR_ARG, NNGP_ERR_COV_MODEL, NNGP_ERR_HIP,
                   NNGP_ERR_INTERRUPT, NNGP_ERR_NOMEM, NNGP_ERR_NOT_PD,
                   NNGP_ERR_UNSUPPORTED, NNGP_FP64, NNGP_MIXED, NNGP_OK,
                   NngpError, exported_symbols, load)
from .nngp import (COV_MODELS, CvGrid, Model, Q, RRng, Server, cNNGP, cov_index, device_count,
                   mkNNIndx, mkNNIndxCB, mkUIndx, n_indx, nn2, prior_draw,
                   rNNGP, rNNGPPredict, rNNGPReplicated, sNNGP, sNNGPPredict,
                   updateBF, updateConjBF)
from .spnngp import (crps, fitted, load_fit, predict, rmspe, save_fit, serve,
                     spConjNNGP, spNNGP)

__all__ = [
    "COV_MODELS", "CvGrid", "Model", "NngpError", "Q", "RRng", "Server", "cNNGP", "cov_index",
    "device_count", "exported_symbols", "load", "mkNNIndx", "mkNNIndxCB",
    "n_indx", "nn2", "prior_draw", "rNNGP", "rNNGPPredict", "rNNGPReplicated",
    "mkUIndx", "sNNGP", "sNNGPPredict", "updateBF",
    "updateConjBF", "spNNGP", "spConjNNGP", "predict", "fitted", "crps",
    "rmspe", "save_fit", "load_fit", "serve",
    "NNGP_OK", "NNGP_ERR_ARG", "NNGP_ERR_NOT_PD", "NNGP_ERR_COV_MODEL",
    "NNGP_ERR_HIP", "NNGP_ERR_NOMEM", "NNGP_ERR_INTERRUPT",
    "NNGP_ERR_UNSUPPORTED", "NNGP_FP64", "NNGP_MIXED",
]
