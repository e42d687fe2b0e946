"""The learned rank model: a small multilayer perceptron that predicts each
box's skeleton count, trained on the GPU by the library's own forward and
backward kernels (rskx_rank_model_*, mlp_kernels.hip) from the exact ranks
of earlier factorizations.

A trained model passed as opts['rank_model'] to rskelf sizes the
Gram-matrix ID's R and T work buffers of every box to a cap of
1.25 k^ + 4 rows (k^ the predicted rank) instead of n, before the level
loop; the exact pivoted ID stays the judge: a box whose rank exceeds its
cap is redone with full buffers (Factor.info()['rank_fallbacks']).  The
reference (core/id.m) has no such model; the factorization's results are
the same with or without it.

Features (8 per box; see include/rskx.h): reference level, points in the
subtree, box extents, point spread and centroid offset, tolerance and the
kernel's wavenumber times the box size.  Target: log2(1 + rank) / 10.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from .rskelf import FlamError, _check

NFEAT = 8


def target(rank):
    """The regression target of a rank."""
    return np.log2(1.0 + np.asarray(rank, dtype=float)) / 10.0


def rank_of(y):
    """The rank a prediction y stands for."""
    return np.exp2(10.0 * np.asarray(y, dtype=float)) - 1.0


class RankModel:
    """Device-resident network (FP64 parameters)."""

    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        _check(L.lib().rskx_rank_model_create(int(seed), ctypes.byref(h)))
        self._h = h.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().rskx_rank_model_free(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @staticmethod
    def nparams():
        return L.lib().rskx_rank_model_nparams()

    def params(self):
        p = np.zeros(self.nparams())
        _check(L.lib().rskx_rank_model_params(self._h, p.ctypes.data, None))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.shape == (self.nparams(),)
        _check(L.lib().rskx_rank_model_params(self._h, None, p.ctypes.data))

    @staticmethod
    def _xt(X, t=None):
        X = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float64)).reshape(-1, NFEAT).cuda()
        if t is None:
            return X, None
        t = torch.as_tensor(np.ascontiguousarray(t, dtype=np.float64)).reshape(-1).cuda()
        if t.shape[0] != X.shape[0]:
            raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:size: one target per sample")
        return X, t

    def grad(self, X, t):
        """(loss, gradient) of the mean squared error over (X, t), computed by
        the device's backward kernel."""
        Xd, td = self._xt(X, t)
        g = torch.empty(self.nparams(), dtype=torch.float64, device="cuda")
        loss = ctypes.c_double()
        _check(L.lib().rskx_set_stream(torch.cuda.current_stream().cuda_stream))
        _check(L.lib().rskx_rank_model_grad(self._h, Xd.data_ptr(), td.data_ptr(), Xd.shape[0],
                                            g.data_ptr(), ctypes.byref(loss)))
        return loss.value, g.cpu().numpy()

    def train(self, X, t, epochs=200, batch=256, lr=3e-3, seed=0):
        """Adam on minibatches of a seeded shuffle of (X, t); returns the
        final mean squared error over all samples."""
        perm = np.random.default_rng(seed).permutation(len(t))
        Xd, td = self._xt(np.asarray(X)[perm], np.asarray(t)[perm])
        loss = ctypes.c_double()
        _check(L.lib().rskx_set_stream(torch.cuda.current_stream().cuda_stream))
        _check(L.lib().rskx_rank_model_train(self._h, Xd.data_ptr(), td.data_ptr(), Xd.shape[0],
                                             int(epochs), int(batch), float(lr),
                                             ctypes.byref(loss)))
        return loss.value

    def predict(self, X):
        Xd, _ = self._xt(X)
        y = torch.empty(Xd.shape[0], dtype=torch.float64, device="cuda")
        _check(L.lib().rskx_set_stream(torch.cuda.current_stream().cuda_stream))
        _check(L.lib().rskx_rank_model_predict(self._h, Xd.data_ptr(), Xd.shape[0], y.data_ptr()))
        return y.cpu().numpy()


def harvest(factors):
    """Training pairs (X, t) from factorizations recorded with
    opts record_boxes=True."""
    Xs, ts = [], []
    for F in factors:
        X, k = F.box_features()
        Xs.append(X)
        ts.append(target(k))
    return np.concatenate(Xs), np.concatenate(ts)
