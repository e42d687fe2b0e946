"""Boundary discretizations of the demo problems and their exact operators.

These build the data of a LaplaceDL2D kernel (points, normals, weights,
diagonal) on the GPU, following the problem set-up of the reference demos:

  circle(N)     unit circle, benchmark sign convention:
                A_ij = 1/2 delta_ij + w (-1/2pi) ((x_i-x_j).n_j)/|x_i-x_j|^2,
                diagonal limit w (-1/4pi), w = 2pi/N
  ie_circle(N)  rskelf/test/ie_circle.m:34,85-103 (A = D w, diag -1/2(1+1/N))
  ellipse(N,a,b) rskelf/test/ie_ellipse.m:28-37,103-123 with semi-axes (a, b)
  sphere(nface) 3D Laplace single layer on a triangulated unit sphere
                (points and proxies after rskelf/test/ie_sphere.m:36-40)
"""
import math

import torch

from .rskelf import HelmholtzCFIE2D, LaplaceDL2D, LaplaceSL3D, ProxyPoints, ProxyRing


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

    def closed_form_matvec(self, X):
        """A X with the unit circle's closed form A = a I + b 1 1': the
        double-layer entries off the diagonal are analytically all equal.
        a and b come from the first column evaluated entry by entry (b the
        mean of its off-diagonal entries); their rounding spread (differences
        of nearby points) bounds how far this is from the entry-wise matrix,
        ~1e-11 in norm at N = 2^20, so it serves as the cheap inner operator
        of the refinement while true residuals use circulant_matvec.  One
        pass over X instead of two FFTs."""
        if getattr(self, "_ab", None) is None:
            c = self.column0()
            b = c[1:].mean()
            dev = (c[1:] - b).abs().sum()
            if not bool(dev <= 1e-8 * c[0].abs()):
                raise ValueError("closed_form_matvec: off-diagonal entries are not constant "
                                 "(not the unit circle); use circulant_matvec")
            self._ab = ((c[0] - b).item(), b.item())
        a, b = self._ab
        vec = X.dim() == 1
        Xt = X.reshape(self.N, -1).to(torch.float64).t()    # (nrhs, N), rows contiguous
        # row sums in two stages (math.gcd(N, 1024) contiguous pieces per row)
        # so the reduction spreads over many blocks even for few columns
        s = Xt.reshape(Xt.shape[0], math.gcd(self.N, 1024), -1).sum(2).sum(1, keepdim=True)
        Yt = torch.add(s.mul_(b), Xt, alpha=a)
        return Yt.reshape(-1) if vec else Yt.t()

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


# ---------------------------------------------------------------- sphere --
_ICO_FACES = [[1, 12, 6], [1, 6, 2], [1, 2, 8], [1, 8, 11], [1, 11, 12], [2, 6, 10], [6, 12, 5],
              [12, 11, 3], [11, 8, 7], [8, 2, 9], [4, 10, 5], [4, 5, 3], [4, 3, 7], [4, 7, 9],
              [4, 9, 10], [5, 10, 6], [3, 5, 12], [7, 3, 11], [9, 7, 8], [10, 9, 2]]


def icosphere(n, by="faces"):
    """Unit-sphere triangulation from a subdivided icosahedron with at least
    n faces (or vertices), in the vertex and face order of
    geom/trisphere_subdiv.m:12-86: new vertices are the projected edge
    midpoints in ascending lexicographic order, appended after the old
    ones; each face is replaced by its four children, listed child-major
    (all first children, then all second, ...).  FP64 tensors on the CPU:
    vertices (nv, 3), faces (nf, 3) 0-based."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = torch.tensor([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                      [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                     dtype=torch.float64)
    V = V / torch.linalg.vector_norm(V[0])
    F = torch.tensor(_ICO_FACES, dtype=torch.int64) - 1
    size = (lambda: V.shape[0]) if by == "vertices" else (lambda: F.shape[0])
    while size() < n:
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        mids = 0.5 * torch.cat([V[a] + V[b], V[a] + V[c], V[b] + V[c]])
        mids = mids / torch.linalg.vector_norm(mids, dim=1, keepdim=True)
        uniq, inv = torch.unique(mids, dim=0, return_inverse=True)
        nv, nf = V.shape[0], F.shape[0]
        ab, ac, bc = (nv + inv[:nf], nv + inv[nf:2 * nf], nv + inv[2 * nf:])
        V = torch.cat([V, uniq])
        F = torch.cat([torch.stack([a, ab, ac], 1), torch.stack([ab, b, bc], 1),
                       torch.stack([ac, bc, c], 1), torch.stack([ab, bc, ac], 1)])
    return V, F


def _self_panel(V, F, C):
    """Integral of 1/(4 pi r) over each flat triangle from its centroid:
    sum over the edges of h (asinh(s2/h) - asinh(s1/h)) / (4 pi), h the
    centroid's distance to the edge's line, s1 < s2 the signed positions
    of the edge's ends along it from the foot of the perpendicular."""
    tot = torch.zeros(F.shape[0], dtype=torch.float64)
    for e in range(3):
        a, b = V[F[:, e]], V[F[:, (e + 1) % 3]]
        edge = b - a
        length = torch.linalg.vector_norm(edge, dim=1)
        u = edge / length[:, None]
        s1 = ((a - C) * u).sum(1)
        h = torch.linalg.vector_norm(a - s1[:, None] * u - C, dim=1)
        tot += h * (torch.asinh((s1 + length) / h) - torch.asinh(s1 / h))
    return tot / (4 * math.pi)


class Surface:
    """Collocation points of a triangulated surface with the 3D Laplace
    single-layer kernel (LaplaceSL3D) and its point proxies."""

    def __init__(self, x, w, diag, scale, proxy):
        self.x, self.w, self.diag, self.scale, self.proxy = x, w, diag, scale, proxy
        self.N = x.shape[1]

    @property
    def kernel(self):
        return LaplaceSL3D(self.w, self.diag, self.scale)


def sphere(nface, nproxy=512, seed=0, device="cuda"):
    """First-kind Laplace single layer on the unit sphere (the C4 workload
    at nface = 327680): triangle centroids of icosphere(nface), A_ij =
    area_j / (4 pi |x_i - x_j|), A_ii the self-panel integral; proxies are
    nproxy vertices of icosphere(nproxy, 'vertices') chosen by a seeded
    permutation (rskelf/test/ie_sphere.m:39-40 draws an unseeded one), at
    one box size from the centre, with incoming rows weighted by the mean
    area 4 pi / N (ie_sphere.m:175-183) and neighbours kept within 1.5 box
    sizes (ie_sphere.m:186)."""
    import numpy as np
    V, F = icosphere(nface)
    P0, P1, P2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    C = (P0 + P1 + P2) / 3
    area = 0.5 * torch.linalg.vector_norm(torch.linalg.cross(P1 - P0, P2 - P1), dim=1)
    diag = _self_panel(V, F, C)
    N = C.shape[0]
    Pv, _ = icosphere(nproxy, by="vertices")
    sel = torch.from_numpy(np.random.default_rng(seed).permutation(Pv.shape[0])[:nproxy])
    proxy = ProxyPoints(Pv[sel].t().contiguous().to(device), radius=1.5, weight=4 * math.pi / N)
    return Surface(C.t().contiguous().to(device), area.to(device), diag.to(device), 1.0, proxy)


# ---------------------------------------------------- Helmholtz CFIE curves --
EULER_GAMMA = 0.57721566490153286061


class HelmholtzCurve:
    """A closed curve with the exterior Dirichlet combined-field matrix
    (HelmholtzCFIE2D) and its combined-field proxy ring."""

    def __init__(self, kind, x, nu, w, diag, k, eta, nproxy):
        self.kind, self.x, self.nu, self.w, self.diag = kind, x, nu, w, diag
        self.k, self.eta = k, eta
        self.N = x.shape[1]
        self.proxy = ProxyRing(p=nproxy, radius=1.5, kind="cfie", weight=float(w.mean()))

    @property
    def kernel(self):
        return HelmholtzCFIE2D(self.nu, self.w, self.diag, self.k, self.eta)

    def matvec(self, X):
        """Exact FP64 complex A X, evaluated on the fly (O(N^2))."""
        from .rskelf import kernel_matvec
        return kernel_matvec(self.kernel, self.x, X)

    def circulant_matvec(self, X):
        """Exact A X on the unit circle, where A is circulant: one complex
        FFT per column with the first column's symbol (as ie_circle.m:47-48
        does for the real matrix); the first column is one product with e_1."""
        if self.kind != "circle":
            raise ValueError("circulant_matvec: the matrix is circulant only on the circle")
        if getattr(self, "_symbol", None) is None:
            e1 = torch.zeros(self.N, dtype=torch.complex128, device=self.x.device)
            e1[0] = 1
            self._symbol = torch.fft.fft(self.matvec(e1))
        vec = X.dim() == 1
        Xt = X.reshape(self.N, -1).to(torch.complex128).t()
        Y = torch.fft.ifft(self._symbol * torch.fft.fft(Xt, dim=1), dim=1).t()
        return Y.reshape(-1) if vec else Y


def helmholtz_curve(N, kind="star", k=200.0, eta=None, nproxy=64, device="cuda"):
    """Exterior Dirichlet Helmholtz combined field A = 1/2 I + D - i eta S on
    the unit circle or the star r(t) = 1 + 0.3 cos 5t, t_i = 2 pi i / N,
    i = 1..N, eta = k by default: arc-length weights h_i; diagonal 1/2
    - h_i kappa_i / 4 pi (the double layer's limit) - i eta s_i with the
    single layer's corrected-trapezoidal self weight s_i = h_i ((i/4) -
    (log(k h_i / 4 pi) + gamma) / 2 pi) (Kapur and Rokhlin's weight
    h log(h / 2 pi) for the log|t| singularity; third-order convergence);
    outward normals."""
    eta = float(k if eta is None else eta)
    t = _theta(N, device)
    if kind == "circle":
        r = torch.ones_like(t)
        rp = torch.zeros_like(t)
        rpp = torch.zeros_like(t)
    elif kind == "star":
        r, rp, rpp = 1 + 0.3 * torch.cos(5 * t), -1.5 * torch.sin(5 * t), -7.5 * torch.cos(5 * t)
    else:
        raise ValueError(kind)
    c, s = torch.cos(t), torch.sin(t)
    x = torch.stack([r * c, r * s])
    xp, yp = rp * c - r * s, rp * s + r * c
    xpp, ypp = rpp * c - 2 * rp * s - r * c, rpp * s + 2 * rp * c - r * s
    speed = torch.sqrt(xp * xp + yp * yp)
    nu = torch.stack([yp, -xp]) / speed
    kappa = (xp * ypp - yp * xpp) / speed ** 3
    h = (2 * math.pi / N) * speed
    re = 0.5 - h * kappa / (4 * math.pi) + eta * h / 4
    im = eta * h * (torch.log(k * h / (4 * math.pi)) + EULER_GAMMA) / (2 * math.pi)
    diag = torch.complex(re, im)
    return HelmholtzCurve(kind, x, nu, h, diag, float(k), eta, nproxy)
