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


# --------------------------------------------------------------------------
# 3D: triangulated unit sphere, Laplace single layer
# --------------------------------------------------------------------------
def trisphere_subdiv(n, vfmin="f"):
    """Vertices (3 x nv) and faces (3 x nf, 0-based) of the unit sphere from
    a recursively subdivided icosahedron, subdividing until the vertex
    ('v') or face ('f') count is at least n; geom/trisphere_subdiv.m:12-86
    (base icosahedron, edge midpoints projected to the sphere, duplicate
    midpoints merged by a sorted unique over rows, each face split in 4)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=float)
    V = V / np.linalg.norm(V[0])
    F = np.array([[1, 12, 6], [1, 6, 2], [1, 2, 8], [1, 8, 11], [1, 11, 12],
                  [2, 6, 10], [6, 12, 5], [12, 11, 3], [11, 8, 7], [8, 2, 9],
                  [4, 10, 5], [4, 5, 3], [4, 3, 7], [4, 7, 9], [4, 9, 10],
                  [5, 10, 6], [3, 5, 12], [7, 3, 11], [9, 7, 8], [10, 9, 2]]) - 1
    count = (lambda: V.shape[0]) if vfmin == "v" else (lambda: F.shape[0])
    while count() < n:
        M = 0.5 * np.concatenate([V[F[:, 0]] + V[F[:, 1]],
                                  V[F[:, 0]] + V[F[:, 2]],
                                  V[F[:, 1]] + V[F[:, 2]]])
        M = M / np.sqrt((M ** 2).sum(axis=1))[:, None]
        M, I = np.unique(M, axis=0, return_inverse=True)
        I = I.reshape(-1)
        nv, nf = V.shape[0], F.shape[0]
        F12, F13, F23 = nv + I[:nf], nv + I[nf:2 * nf], nv + I[2 * nf:]
        V = np.concatenate([V, M])
        F = np.concatenate([np.stack([F[:, 0], F12, F13], 1), np.stack([F12, F[:, 1], F23], 1),
                            np.stack([F13, F23, F[:, 2]], 1), np.stack([F12, F23, F13], 1)])
    return V.T.copy(), F.T.copy()


def tri3geom(V, F):
    """Centroid, unit normal and area of each triangle (geom/tri3geom.m:9-22)."""
    P0, P1, P2 = V[:, F[0]], V[:, F[1]], V[:, F[2]]
    C = (P0 + P1 + P2) / 3.0
    Nv = np.cross((P1 - P0).T, (P2 - P1).T).T
    A = np.sqrt((Nv ** 2).sum(axis=0))
    return C, Nv / A, 0.5 * A


def self_panel_sl(V, F, C):
    """Integral of 1/(4 pi |c - y|) over each flat triangle for a point c in
    its plane and inside it (the centroid): split the triangle at c into
    three triangles with apex c; over one of them, with its base edge at
    distance h from c and the base's ends at signed offsets s1 < s2 from
    the foot of the perpendicular, the integral of 1/r in polar
    coordinates about c is h (asinh(s2/h) - asinh(s1/h))."""
    out = np.zeros(F.shape[1])
    for e in range(3):
        a = V[:, F[e]]
        b = V[:, F[(e + 1) % 3]]
        u = b - a
        L = np.sqrt((u ** 2).sum(axis=0))
        u = u / L
        s1 = ((a - C) * u).sum(axis=0)            # offsets of the ends along the edge
        s2 = s1 + L
        foot = a - s1 * u
        h = np.sqrt(((foot - C) ** 2).sum(axis=0))
        out += h * (np.arcsinh(s2 / h) - np.arcsinh(s1 / h))
    return out / (4.0 * math.pi)


def self_panel_sl_quadrature(V, F, C, order=24):
    """The same integral by brute force: Duffy-mapped Gauss-Legendre
    quadrature on the three sub-triangles with apex at c (the map's
    Jacobian cancels the 1/r singularity), the construction of
    quad/quad_sqtri3.m.  A check of self_panel_sl, not used otherwise."""
    g, wg = np.polynomial.legendre.leggauss(order)
    g = 0.5 * (g + 1.0)
    wg = 0.5 * wg
    out = np.zeros(F.shape[1])
    for e in range(3):
        a = V[:, F[e]]
        b = V[:, F[(e + 1) % 3]]
        for s, ws in zip(g, wg):
            for t, wt in zip(g, wg):
                # y = c + s ((a - c) + t (b - a)), dA = s |(a-c) x (b-a)| ds dt
                y = C + s * ((a - C) + t * (b - a))
                jac = s * np.linalg.norm(np.cross((a - C).T, (b - a).T), axis=1)
                r = np.sqrt(((y - C) ** 2).sum(axis=0))
                out += ws * wt * jac / r
    return out / (4.0 * math.pi)


class SphereSL:
    """First-kind Laplace single-layer matrix on the triangulated unit
    sphere: collocation at triangle centroids with one-point quadrature,
    A_ij = area_j / (4 pi |x_i - x_j|) for i != j, and A_ii the exact
    integral of the kernel over triangle i at its centroid.  Points and
    areas from trisphere_subdiv / tri3geom (rskelf/test/ie_sphere.m:36-37).
    Proxies: nproxy vertices of trisphere_subdiv(nproxy, 'v'), picked by a
    seeded permutation (ie_sphere.m:39-40 uses an unseeded randperm), at
    pxy = proxy l + ctr (ie_sphere.m:175); rows for the outgoing (area-
    weighted) and, for symm 'n', the incoming (4 pi / N weighted)
    interactions (ie_sphere.m:178-183); neighbour filter radius 1.5
    (ie_sphere.m:186)."""

    def __init__(self, nface, nproxy=512, seed=0, symm="n"):
        self.V, self.F = trisphere_subdiv(nface)
        self.x, self.nu, self.area = tri3geom(self.V, self.F)
        self.N = self.x.shape[1]
        self.diag = self_panel_sl(self.V, self.F, self.x)
        Pv, _ = trisphere_subdiv(nproxy, "v")
        sel = np.random.default_rng(seed).permutation(Pv.shape[1])[:nproxy]
        self.proxy = Pv[:, sel]
        self.two_sided = symm == "n"

    @classmethod
    def from_arrays(cls, x, area, diag, proxy, symm="n"):
        """The same problem on given points, weights, diagonal and proxies
        (to run the oracle on exactly the arrays another build produced)."""
        self = cls.__new__(cls)
        self.x, self.area, self.diag, self.proxy = (np.asarray(a, dtype=float)
                                                    for a in (x, area, diag, proxy))
        self.N = self.x.shape[1]
        self.two_sided = symm == "n"
        return self

    @staticmethod
    def sl(xt, ys):
        d = xt[:, :, None] - ys[:, None, :]
        with np.errstate(divide="ignore"):
            return 1.0 / (4.0 * math.pi * np.sqrt((d ** 2).sum(axis=0)))

    def afun(self, I, J):
        I = np.asarray(I, dtype=np.int64)
        J = np.asarray(J, dtype=np.int64)
        A = self.sl(self.x[:, I], self.x[:, J]) * self.area[J][None, :]
        same = I[:, None] == J[None, :]
        return np.where(same, self.diag[J][None, :], A)

    def pxyfun(self, x, slf, nbr, l, ctr):
        pxy = self.proxy * l[:, None] + ctr[:, None]
        Kout = self.sl(pxy, x[:, slf]) * self.area[slf][None, :]
        if self.two_sided:
            Kin = self.sl(x[:, slf], pxy).T * (4.0 * math.pi / self.N)
            Kpxy = np.concatenate([Kout, Kin])
        else:
            Kpxy = Kout
        keep = np.sum(((x[:, nbr] - ctr[:, None]) / l[:, None]) ** 2, axis=0) < 1.5 ** 2
        return Kpxy, nbr[keep]

    def dense(self):
        idx = np.arange(self.N)
        return self.afun(idx, idx)

    def matvec(self, X):
        return self.dense() @ np.asarray(X)


# --------------------------------------------------------------------------
# PDE known-answer checks (the demos' end-to-end accuracy tests)
# --------------------------------------------------------------------------
def laplace2d_pde_error(x, nu, h, solve, ratio=1.0, seed=0):
    """Interior Dirichlet problem with a known harmonic solution, as
    rskelf/test/ie_circle.m:65-81 and ie_ellipse.m:83-99 run it: 16 point
    charges on the curve (ratio cos t, sin t) scaled by 2 (outside) make
    the boundary data B = G(x, src) q with G = -log r / 2pi; the density
    X = solve(B) of the double-layer equation gives the field at 16 targets
    scaled by 0.5 (inside) as (dG/dn_y w) X, compared with G(trg, src) q.
    The charges are seeded here (the demos draw an unseeded rand).
    Returns the relative error of the 16 field values."""
    m = 16
    th = np.arange(1, m + 1) * 2.0 * math.pi / m
    src = 2.0 * np.stack([ratio * np.cos(th), np.sin(th)])
    trg = 0.5 * np.stack([ratio * np.cos(th), np.sin(th)])
    q = np.random.default_rng(seed).random(m)

    def g(a, b):
        dx = a[0][:, None] - b[0][None, :]
        dy = a[1][:, None] - b[1][None, :]
        return -np.log(np.sqrt(dx * dx + dy * dy)) / (2.0 * math.pi)

    B = g(x, src) @ q
    X = np.asarray(solve(B))
    Y = (CurveDL.dl(trg, x, nu) * h[None, :]) @ X
    Z = g(trg, src) @ q
    return float(np.linalg.norm(Z - Y) / np.linalg.norm(Z))


def sphere_pde_error(x, area, solve, seed=0):
    """The known-answer check of rskelf/test/ie_sphere.m:127-145 adapted to
    the first-kind single layer: 16 seeded unit charges on the sphere of
    radius 2 (outside) give boundary data B = G(x, src) q with
    G = 1 / (4 pi r); the density X = solve(B) of S X = B gives the field
    at 16 seeded points on the sphere of radius 0.5 (inside) as
    (G(trg, x) area) X, compared with G(trg, src) q.  One-point quadrature
    makes this converge at first order in the mesh width."""
    rng = np.random.default_rng(seed)
    src = rng.standard_normal((3, 16))
    src = 2.0 * src / np.linalg.norm(src, axis=0)
    trg = rng.standard_normal((3, 16))
    trg = 0.5 * trg / np.linalg.norm(trg, axis=0)
    q = rng.random(16)
    B = SphereSL.sl(x, src) @ q
    X = np.asarray(solve(B))
    Y = (SphereSL.sl(trg, x) * area[None, :]) @ X
    Z = SphereSL.sl(trg, src) @ q
    return float(np.linalg.norm(Z - Y) / np.linalg.norm(Z))
