"""Problem definitions for the oracle (test infrastructure).

Matrix-entry and proxy callbacks with the signatures rskelf expects
(Afun(I, J), pxyfun(x, slf, nbr, l, ctr)), following the reference demos:

  CurveDL('circle')     the unit-circle Laplace double-layer matrix written
                        in the benchmark definition: A_ij = 1/2 delta_ij +
                        w (-1/2pi) ((x_i-x_j).n_j)/|x_i-x_j|^2, diagonal
                        limit w (-1/4pi); single-layer proxy ring as in
                        rskelf/test/ie_circle.m:35,106-115.
  CurveDL('ie_circle')  the demo's own matrix, rskelf/test/ie_circle.m:85-103
                        (A = D w, diagonal -1/2 (1 + 1/N)).
  CurveDL('ellipse')    rskelf/test/ie_ellipse.m:28-37,103-132 with semi-axes
                        (a, b) and double-layer proxies.
"""
import math

import numpy as np


class CurveDL:
    """Laplace double-layer Nystrom matrix on a closed 2D curve."""

    def __init__(self, N, kind="circle", nproxy=64, a=1.0, b=1.0):
        self.N = N
        self.kind = kind
        th = np.arange(1, N + 1) * 2.0 * math.pi / N
        if kind in ("circle", "ie_circle"):
            self.x = np.stack([np.cos(th), np.sin(th)])
            self.nu = self.x.copy()
            self.h = np.full(N, 2.0 * math.pi / N)
            self.kappa = np.ones(N)
        elif kind == "ellipse":                     # ie_ellipse.m:28-35
            self.x = np.stack([a * np.cos(th), b * np.sin(th)])
            nu = np.stack([b * np.cos(th), a * np.sin(th)])
            sp = np.sqrt(nu[0] ** 2 + nu[1] ** 2)
            self.nu = nu / sp
            self.kappa = a * b / sp ** 3
            self.h = 2.0 * math.pi / N * sp
        else:
            raise ValueError(kind)
        tp = np.arange(1, nproxy + 1) * 2.0 * math.pi / nproxy
        self.proxy = 1.5 * np.stack([np.cos(tp), np.sin(tp)])   # ie_circle.m:35

    # -- kernels -----------------------------------------------------------
    @staticmethod
    def dl(xt, ys, nus):
        """dG/dn_y with G = -log(r)/2pi, as the demos write it:
        ((x-y).n_y) / (2 pi r^2)  (ie_circle.m:88-97)."""
        dx = xt[0][:, None] - ys[0][None, :]
        dy = xt[1][:, None] - ys[1][None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            return (dx * nus[0][None, :] + dy * nus[1][None, :]) / (2.0 * math.pi * (dx * dx + dy * dy))

    @staticmethod
    def sl(xt, ys):
        dx = xt[0][:, None] - ys[0][None, :]
        dy = xt[1][:, None] - ys[1][None, :]
        return -np.log(np.sqrt(dx * dx + dy * dy)) / (2.0 * math.pi)

    # -- callbacks ---------------------------------------------------------
    def afun(self, I, J):
        I = np.asarray(I, dtype=np.int64)
        J = np.asarray(J, dtype=np.int64)
        D = self.dl(self.x[:, I], self.x[:, J], self.nu[:, J]) * self.h[J][None, :]
        same = I[:, None] == J[None, :]
        if self.kind == "circle":
            D = -D
            diag = 0.5 - self.h[J] / (4.0 * math.pi)
        elif self.kind == "ie_circle":                # ie_circle.m:102
            diag = np.full(len(J), -0.5 * (1.0 + 1.0 / self.N))
        else:                                         # ie_ellipse.m:120-122
            diag = -0.5 * (1.0 + self.h[J] * self.kappa[J] / (2.0 * math.pi))
        D = np.where(same, diag[None, :], D)
        return D

    def pxyfun(self, x, slf, nbr, l, ctr):
        """Proxy block and neighbour filter (ie_circle.m:106-115,
        ie_ellipse.m:126-132)."""
        pxy = self.proxy * l[:, None] + ctr[:, None]
        if self.kind == "ellipse":
            Kpxy = self.dl(pxy, x[:, slf], self.nu[:, slf]) * self.h[slf][None, :]
        else:
            Kpxy = self.sl(pxy, x[:, slf]) * (2.0 * math.pi / self.N)
        keep = np.sum(((x[:, nbr] - ctr[:, None]) / l[:, None]) ** 2, axis=0) < 1.5 ** 2
        return Kpxy, nbr[keep]

    # -- exact operators ---------------------------------------------------
    def dense(self):
        idx = np.arange(self.N)
        return self.afun(idx, idx)

    def circulant_column(self):
        """First column of the (circulant) circle matrix by direct kernel
        evaluation, as ie_circle.m:47 builds it."""
        return self.afun(np.arange(self.N), np.arange(1))[:, 0]

    def matvec(self, X):
        """Exact FP64 product: FFT circulant for the circle (ie_circle.m:47-48,
        118-120), dense rows otherwise."""
        X = np.asarray(X)
        if self.kind in ("circle", "ie_circle"):
            G = np.fft.fft(self.circulant_column())
            Y = np.fft.ifft(G[:, None] * np.fft.fft(X.reshape(self.N, -1), axis=0), axis=0)
            Y = Y.real if not np.iscomplexobj(X) else Y
            return Y.reshape(X.shape)
        return self.dense() @ X

    def closed_form_inverse_apply(self, B):
        """A = alpha I + beta 1 1^T on the unit circle (the double-layer
        off-diagonal is exactly constant there), so
        A^{-1} = (I - beta/(alpha + N beta) 1 1^T) / alpha."""
        N = self.N
        if self.kind == "circle":
            alpha, beta = 0.5 - 1.0 / N, 0.5 / N
        elif self.kind == "ie_circle":
            alpha, beta = -0.5, -0.5 / N
        else:
            raise ValueError("closed form only for the circle")
        B = np.asarray(B, dtype=float)
        s = B.sum(axis=0)
        return (B - beta / (alpha + N * beta) * s) / alpha
