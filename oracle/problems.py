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


# --------------------------------------------------------------------------
# 2D Helmholtz combined-field integral equation (complex)
# --------------------------------------------------------------------------
EULER_GAMMA = 0.57721566490153286061


class CurveCFIE:
    """Exterior Dirichlet Helmholtz problem on a closed 2D curve by the
    combined-field integral equation (Brakhage-Werner): u = D[s] - i eta S[s]
    with G(x, y) = (i/4) H0^(1)(k |x - y|), whose exterior trace gives
        A = 1/2 I + D - i eta S,
        A_ij = h_j (dG/dn_j(x_i, x_j) - i eta G(x_i, x_j)),  i != j,
    dG/dn_y(x, y) = (i k / 4) H1^(1)(k r) ((x - y) . n_y) / r, on the
    trapezoidal rule in the curve parameter t_i = 2 pi i / N, i = 1..N, with
    arc-length weights h_j.  The kernel follows rskelf/test/ie_square3.m:44,
    107-110 (Kfun = 0.25i besselh(0, 1, k r)); the combined layer and the
    curves are the builder's (the reference has no curve CFIE demo).

    Diagonal: 1/2 (jump) + h_i (-kappa_i / 4 pi) (the double layer's
    continuous limit, as ie_ellipse.m:120-122 takes it) - i eta s_i with
    s_i = h_i ((i/4) - (log(k h_i / 4 pi) + gamma) / 2 pi): the single layer's
    smooth part at r = 0 plus the corrected-trapezoidal weight h log(h / 2 pi)
    of its log|t| singularity (the zeta'(0) term of the Euler-Maclaurin
    expansion for a punctured periodic trapezoidal rule; Kapur and Rokhlin,
    SIAM J. Numer. Anal. 34 (1997)), which makes the rule converge at third
    order in h.

    kind 'circle': the unit circle (then A is circulant and complex
    symmetric); 'star': r(t) = 1 + 0.3 cos 5t.

    Proxies (pxyfun): a ring at 1.5 box sizes (as ie_circle.m:35,107) of
    p_l = max(p, 8 ceil((1.5 k l + 48) / 8)) points, l the box's longest
    side: a box's outgoing field on the ring has angular bandwidth about
    k l / sqrt 2 + O(log 1/tol), so the count grows with k l.  Rows: the
    outgoing interactions A(pxy, slf) (the matrix kernel with the proxy
    point as target) and, for symm 'n', the incoming ones as the conjugate
    of a combined-field source on the ring, w (dG/dn_q - i eta G)(x_j, y_q)
    with the ring's outward normal n_q and w the mean weight (the interior
    field of far sources is a combined-field potential on the ring for
    every k, so no interior resonance of the ring can drop a direction);
    neighbours kept within 1.5 box sizes (ie_circle.m:114)."""

    def __init__(self, N, kind="star", k=20.0, eta=None, nproxy=64, symm="n"):
        self.N, self.kind, self.k = N, kind, float(k)
        self.eta = float(k if eta is None else eta)
        self.nproxy = nproxy
        self.two_sided = symm == "n"
        t = np.arange(1, N + 1) * 2.0 * math.pi / N
        if kind == "circle":
            r, rp, rpp = np.ones(N), np.zeros(N), np.zeros(N)
        elif kind == "star":
            r, rp, rpp = 1 + 0.3 * np.cos(5 * t), -1.5 * np.sin(5 * t), -7.5 * np.cos(5 * t)
        else:
            raise ValueError(kind)
        c, s = np.cos(t), np.sin(t)
        self.x = np.stack([r * c, r * s])
        xp, yp = rp * c - r * s, rp * s + r * c
        xpp, ypp = rpp * c - 2 * rp * s - r * c, rpp * s + 2 * rp * c - r * s
        speed = np.sqrt(xp ** 2 + yp ** 2)
        self.nu = np.stack([yp, -xp]) / speed
        self.kappa = (xp * ypp - yp * xpp) / speed ** 3
        self.h = 2.0 * math.pi / N * speed
        self.pw = float(self.h.mean())
        h, kk = self.h, self.k
        ssl = h * (0.25j - (np.log(kk * h / (4.0 * math.pi)) + EULER_GAMMA) / (2.0 * math.pi))
        self.diag = 0.5 - h * self.kappa / (4.0 * math.pi) - 1j * self.eta * ssl

    @classmethod
    def from_arrays(cls, x, nu, h, diag, k, eta, pw, nproxy=64, symm="n"):
        """The same problem on given points, normals, weights and diagonal."""
        self = cls.__new__(cls)
        self.x, self.nu, self.h = (np.asarray(a, dtype=float) for a in (x, nu, h))
        self.diag = np.asarray(diag, dtype=complex)
        self.N = self.x.shape[1]
        self.k, self.eta, self.pw, self.nproxy = float(k), float(eta), float(pw), nproxy
        self.two_sided = symm == "n"
        self.kind = "given"
        return self

    # -- kernels -----------------------------------------------------------
    def g(self, xt, ys):
        from scipy.special import hankel1
        dx = xt[0][:, None] - ys[0][None, :]
        dy = xt[1][:, None] - ys[1][None, :]
        with np.errstate(all="ignore"):
            return 0.25j * hankel1(0, self.k * np.sqrt(dx * dx + dy * dy))

    def dg(self, xt, ys, nus):
        from scipy.special import hankel1
        dx = xt[0][:, None] - ys[0][None, :]
        dy = xt[1][:, None] - ys[1][None, :]
        r = np.sqrt(dx * dx + dy * dy)
        with np.errstate(all="ignore"):
            return (0.25j * self.k * hankel1(1, self.k * r) *
                    (dx * nus[0][None, :] + dy * nus[1][None, :]) / r)

    def cfie(self, xt, ys, nus, ws):
        """h_j (dG/dn_j - i eta G)(x_t, y_j): the matrix kernel."""
        return (self.dg(xt, ys, nus) - 1j * self.eta * self.g(xt, ys)) * ws[None, :]

    # -- callbacks ---------------------------------------------------------
    def afun(self, I, J):
        I = np.asarray(I, dtype=np.int64)
        J = np.asarray(J, dtype=np.int64)
        A = self.cfie(self.x[:, I], self.x[:, J], self.nu[:, J], self.h[J])
        same = I[:, None] == J[None, :]
        return np.where(same, self.diag[J][None, :], A)

    def nproxy_at(self, l):
        return max(self.nproxy, 8 * int(math.ceil((1.5 * self.k * float(np.max(l)) + 48) / 8)))

    def proxy_ring(self, l, ctr):
        p = self.nproxy_at(l)
        ph = np.arange(1, p + 1) * 2.0 * math.pi / p
        unit = 1.5 * np.stack([np.cos(ph), np.sin(ph)])
        pxy = unit * l[:, None] + ctr[:, None]
        nq = unit / l[:, None]                    # normal of the (possibly elliptic) ring
        nq = nq / np.sqrt((nq ** 2).sum(axis=0))
        return pxy, nq

    def pxyfun(self, x, slf, nbr, l, ctr):
        pxy, nq = self.proxy_ring(l, ctr)
        Kout = self.cfie(pxy, x[:, slf], self.nu[:, slf], self.h[slf])
        if self.two_sided:
            Kin = self.cfie(x[:, slf], pxy, nq, np.full(pxy.shape[1], self.pw)).T.conj()
            Kpxy = np.concatenate([Kout, Kin])
        else:
            Kpxy = Kout
        keep = np.sum(((x[:, nbr] - ctr[:, None]) / l[:, None]) ** 2, axis=0) < 1.5 ** 2
        return Kpxy, nbr[keep]

    # -- exact operators ---------------------------------------------------
    def dense(self):
        idx = np.arange(self.N)
        return self.afun(idx, idx)

    def matvec(self, X):
        X = np.asarray(X)
        if self.kind == "circle":                 # circulant (ie_circle.m:47-48 pattern)
            col = self.afun(np.arange(self.N), np.arange(1))[:, 0]
            Y = np.fft.ifft(np.fft.fft(col)[:, None] * np.fft.fft(X.reshape(self.N, -1), axis=0),
                            axis=0)
            return Y.reshape(X.shape)
        return self.dense() @ X

    def circle_eigenvalue(self, m):
        """lambda_m of the continuous exterior-trace operator on the unit
        circle, A e^{imt} = lambda_m e^{imt}: with the addition theorem
        G = (i/4) sum_n J_n(k|y|) H_n(k|x|) e^{in(t_x - t_y)} (|x| > |y|),
        D e^{imt} -> (i pi k / 2) J_m'(k) H_m(k) and S e^{imt} ->
        (i pi / 2) J_m(k) H_m(k) on the circle, so
        lambda_m = (i pi k / 2) J_m'(k) H_m(k) + (eta pi / 2) J_m(k) H_m(k)."""
        from scipy.special import hankel1, jv, jvp
        k = self.k
        return ((0.5j * math.pi * k) * jvp(m, k) * hankel1(m, k) +
                (0.5 * self.eta * math.pi) * jv(m, k) * hankel1(m, k))


def helmholtz_pde_error(pr, solve, x0=(0.1, 0.2), radius=2.5):
    """Exterior Dirichlet known answer: the radiating field G(x, x0) of a
    source inside the curve is the boundary data; the density X = solve(B)
    gives the field at 16 points on a circle of `radius` (outside) as the
    combined-field potential, compared with G(trg, x0) (the known-answer
    pattern of rskelf/test/ie_circle.m:65-81 for this exterior problem)."""
    src = np.asarray(x0, dtype=float).reshape(2, 1)
    B = pr.g(pr.x, src)[:, 0]
    X = np.asarray(solve(B))
    ang = np.arange(16) * 2.0 * math.pi / 16
    trg = radius * np.stack([np.cos(ang), np.sin(ang)])
    Y = pr.cfie(trg, pr.x, pr.nu, pr.h) @ X
    Z = pr.g(trg, src)[:, 0]
    return float(np.linalg.norm(Y - Z) / np.linalg.norm(Z))
