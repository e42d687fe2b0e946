"""Boundary discretizations of the demo problems and their exact operators.

These build the data of a LaplaceDL2D kernel (points, normals, weights,
diagonal) on the GPU, following the problem set-up of the reference demos:

  circle(N)     unit circle, benchmark sign convention:
                A_ij = 1/2 delta_ij + w (-1/2pi) ((x_i-x_j).n_j)/|x_i-x_j|^2,
                diagonal limit w (-1/4pi), w = 2pi/N
  ie_circle(N)  rskelf/test/ie_circle.m:34,85-103 (A = D w, diag -1/2(1+1/N))
  ellipse(N,a,b) rskelf/test/ie_ellipse.m:28-37,103-123 with semi-axes (a, b)
"""
import math

import torch

from .rskelf import LaplaceDL2D, ProxyRing


class Curve:
    """A closed curve discretized by the trapezoidal rule."""

    def __init__(self, x, nu, w, diag, scale, proxy):
        self.x, self.nu, self.w, self.diag = x, nu, w, diag
        self.scale = scale
        self.proxy = proxy
        self.N = x.shape[1]

    @property
    def kernel(self):
        return LaplaceDL2D(self.nu, self.w, self.diag, self.scale)

    def column0(self):
        """First column A(:,1), evaluated entry by entry in FP64."""
        x, nu, w = self.x, self.nu, self.w
        dx = x[0] - x[0, 0]
        dy = x[1] - x[1, 0]
        r2 = dx * dx + dy * dy
        c = self.scale * w[0] * (dx * nu[0, 0] + dy * nu[1, 0]) / (2 * math.pi * r2)
        c[0] = self.diag[0]
        return c

    def circulant_matvec(self, X):
        """Exact FP64 A X for rotation-invariant curves (the circle): A is
        circulant, so A X = ifft(fft(A(:,1)) .* fft(X)) (ie_circle.m:47-48),
        here with real transforms along each column and the symbol cached."""
        if getattr(self, "_symbol", None) is None:
            self._symbol = torch.fft.rfft(self.column0())
        vec = X.dim() == 1
        X2 = X.reshape(self.N, -1).to(torch.float64)
        Xt = X2.t()                        # one transform per column, contiguous rows
        Y = torch.fft.irfft(self._symbol * torch.fft.rfft(Xt, dim=1), n=self.N, dim=1).t()
        return Y.reshape(-1) if vec else Y


def _theta(N, device):
    return torch.arange(1, N + 1, dtype=torch.float64, device=device) * (2 * math.pi / N)


def circle(N, nproxy=64, device="cuda"):
    th = _theta(N, device)
    x = torch.stack([torch.cos(th), torch.sin(th)])
    w = torch.full((N,), 2 * math.pi / N, dtype=torch.float64, device=device)
    diag = 0.5 - w / (4 * math.pi)
    proxy = ProxyRing(p=nproxy, radius=1.5, kind="sl", weight=2 * math.pi / N)
    return Curve(x, x.clone(), w, diag, -1.0, proxy)


def ie_circle(N, nproxy=64, device="cuda"):
    th = _theta(N, device)
    x = torch.stack([torch.cos(th), torch.sin(th)])
    w = torch.full((N,), 2 * math.pi / N, dtype=torch.float64, device=device)
    diag = torch.full((N,), -0.5 * (1 + 1 / N), dtype=torch.float64, device=device)
    proxy = ProxyRing(p=nproxy, radius=1.5, kind="sl", weight=2 * math.pi / N)
    return Curve(x, x.clone(), w, diag, 1.0, proxy)


def ellipse(N, a=1.0, b=1.0, nproxy=64, device="cuda"):
    th = _theta(N, device)
    x = torch.stack([a * torch.cos(th), b * torch.sin(th)])
    nu = torch.stack([b * torch.cos(th), a * torch.sin(th)])
    sp = torch.sqrt(nu[0] ** 2 + nu[1] ** 2)
    nu = nu / sp
    kappa = a * b / sp ** 3
    h = 2 * math.pi / N * sp
    diag = -0.5 * (1 + h * kappa / (2 * math.pi))
    proxy = ProxyRing(p=nproxy, radius=1.5, kind="dl", weight=0.0)
    return Curve(x, nu, h, diag, 1.0, proxy)
