"""Pins the CPU oracle (oracle/flam.py) against the reference's published
numbers, closed forms, dense LAPACK and invariants.  CPU only."""
import json
import math
import os
import warnings

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import flam, problems

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _readme():
    with open(os.path.join(GOLDEN, "readme_ie_circle.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- tree -----
def test_hypoct_readme_block_counts():
    """README.md:99-106: the nblk column depends on geometry only."""
    g = _readme()
    pr = problems.CurveDL(g["params"]["N"], "ie_circle")
    t = flam.hypoct(pr.x, g["params"]["occ"])
    assert t.nlvl == 8
    nonempty = np.array([len(v) > 0 for v in t.xi])
    pblk = np.concatenate([[0], np.cumsum([nonempty[t.lvp[l]:t.lvp[l + 1]].sum()
                                           for l in range(t.nlvl)])])
    nblk = [int(pblk[lv - 1] + t.lvp[lv] - t.lvp[lv - 1]) for lv in range(t.nlvl, 0, -1)]
    assert nblk == [r[1] for r in g["rows"]]


def _brute_adjacent(t, i, j):
    """Closed boxes touch (share a face, edge or corner) without overlapping."""
    li, lj = t.l[:, t.level_of(i)], t.l[:, t.level_of(j)]
    gap = np.abs(t.ctr[i] - t.ctr[j]) - 0.5 * (li + lj)
    return bool(np.all(gap <= 1e-12 * li.max()))


@pytest.mark.parametrize("d,n,occ", [(2, 3000, 20), (3, 2000, 40), (1, 500, 7)])
def test_hypoct_invariants(d, n, occ):
    rng = np.random.default_rng(1)
    x = rng.random((d, n)) ** 2        # non-uniform so leaves land on several levels
    t = flam.hypoct(x, occ)
    leaves = [b for b in range(t.nbox) if not t.chld[b]]
    pts = np.concatenate([t.xi[b] for b in leaves])
    assert np.array_equal(np.sort(pts), np.arange(n))            # leaves partition
    for b in leaves:
        assert len(t.xi[b]) <= occ or t.level_of(b) == t.nlvl - 1
        lo = t.ctr[b] - 0.5 * t.l[:, t.level_of(b)]
        hi = t.ctr[b] + 0.5 * t.l[:, t.level_of(b)]
        xb = x[:, t.xi[b]]
        assert np.all(xb >= lo[:, None] - 1e-12) and np.all(xb <= hi[:, None] + 1e-12)
        assert np.all(np.diff(t.xi[b]) > 0)                      # order kept
    for b in range(t.nbox):
        if t.chld[b]:
            assert len(t.xi[b]) == 0
        lv = t.level_of(b)
        for j in t.nbor[b]:
            assert j != b and _brute_adjacent(t, b, j)
            if t.level_of(j) == lv:
                assert b in t.nbor[j]                             # same-level symmetry
            else:
                assert t.level_of(j) < lv and not t.chld[j]       # coarser leaf
        # brute force: every adjacent same-level box is listed
        for j in range(t.lvp[lv], t.lvp[lv + 1]):
            if j != b and _brute_adjacent(t, b, j):
                assert j in t.nbor[b]


def test_hypoct_errors():
    with pytest.raises(flam.FlamError) as e:
        flam.hypoct(np.zeros((2, 3)), -1)
    assert e.value.ident == "FLAM:hypoct:invalidOcc"
    with pytest.raises(flam.FlamError) as e:
        flam.hypoct(np.zeros((2, 3)), 2, lvlmax=0)
    assert e.value.ident == "FLAM:hypoct:invalidLvlmax"


def test_mround_matches_matlab():
    assert list(flam.mround([0.5, 1.5, 2.5, -0.5, -2.5, 0.49])) == [1, 2, 3, -1, -3, 0]


# ------------------------------------------------------------------ id -----
def _lowrank(m, n, k, rng, decay=1e-3):
    U = np.linalg.qr(rng.standard_normal((m, k)))[0]
    V = np.linalg.qr(rng.standard_normal((n, k)))[0]
    s = decay ** (np.arange(k) / max(k - 1, 1))
    return (U * s) @ V.T


@pytest.mark.parametrize("tol", [1e-3, 1e-6, 1e-10])
def test_id_rank_matches_textbook_qrcp(tol):
    """Without refinement the rank is #{|R_jj| > tol |R_11|} of QRCP."""
    rng = np.random.default_rng(2)
    A = _lowrank(90, 60, 40, rng, decay=1e-12)
    sk, rd, T, _ = flam.interp_decomp(A, tol, Tmax=np.inf)
    R = sla.qr(A, mode="r", pivoting=True)[0]
    d = np.abs(np.diag(R))
    assert len(sk) == int(np.sum(d > tol * d[0]))
    assert sorted(np.concatenate([sk, rd]).tolist()) == list(range(60))


@pytest.mark.parametrize("m,n", [(80, 50), (50, 80), (900, 40)])
def test_id_error_bound_and_T_bound(m, n):
    rng = np.random.default_rng(3)
    A = _lowrank(m, n, min(m, n) // 2, rng, decay=1e-9)
    tol = 1e-7
    sk, rd, T, _ = flam.interp_decomp(A, tol)
    err = np.linalg.norm(A[:, rd] - A[:, sk] @ T, 2)
    smax = np.linalg.norm(A, 2)
    assert err <= 100 * tol * smax
    assert np.abs(T).max() <= 2.0 + 1e-12                         # Tmax after RRQR
    # SVD brute force: a rank below the tol-numerical rank cannot reach tol
    s = np.linalg.svd(A, compute_uv=False)
    assert len(sk) >= int(np.sum(s > 1.01 * err)) - 1


def _kahan(n, c):
    """Kahan's matrix: column pivoting keeps the natural order, so plain
    QRCP gives huge interpolation coefficients."""
    s = math.sqrt(1 - c * c)
    K = np.diag(s ** np.arange(n)) @ (np.triu(-c * np.ones((n, n)), 1) + np.eye(n))
    return K @ np.diag(1 - 1e-10 * np.arange(n))


def test_id_refinement_swaps_large_entries():
    A = _kahan(50, 0.7)
    sk0, rd0, T0, _ = flam.interp_decomp(A, 1e-6, Tmax=np.inf)
    assert np.abs(T0).max() > 1e4
    sk, rd, T, niter = flam.interp_decomp(A, 1e-6, Tmax=2)
    assert niter >= 1 and len(sk) == len(sk0)
    assert np.abs(T).max() <= 2 + 1e-10
    assert np.linalg.norm(A[:, rd] - A[:, sk] @ T, 2) <= 1e-4 * np.linalg.norm(A, 2)


def test_id_special_cases():
    sk, rd, T, _ = flam.interp_decomp(np.zeros((0, 5)), 1e-6)
    assert len(sk) == 0 and list(rd) == list(range(5)) and T.shape == (0, 5)
    sk, rd, T, _ = flam.interp_decomp(np.zeros((4, 0)), 1e-6)
    assert len(sk) == 0 and len(rd) == 0
    rng = np.random.default_rng(5)
    A = rng.standard_normal((30, 12))
    sk, rd, T, _ = flam.interp_decomp(A, 5)                          # fixed rank 5
    assert len(sk) == 5 and T.shape == (5, 7)
    sk, rd, T, _ = flam.interp_decomp(A, 0)                          # rank 0 => kmax = n
    assert len(sk) == 12
    v = rng.standard_normal((7, 1))
    sk, rd, T, _ = flam.interp_decomp(v @ rng.standard_normal((1, 9)), 1e-12)
    assert len(sk) == 1
    with pytest.raises(flam.FlamError) as e:
        flam.interp_decomp(A, -1)
    assert e.value.ident == "FLAM:id:invalidRankOrTol"
    with pytest.raises(flam.FlamError) as e:
        flam.interp_decomp(A, 1e-3, Tmax=0.5)
    assert e.value.ident == "FLAM:id:invalidTmax"
    with pytest.raises(flam.FlamError) as e:
        flam.interp_decomp(A, 1e-3, rrqr_iter=-1)
    assert e.value.ident == "FLAM:id:invalidRRQRIter"


def test_id_iteration_cap_warns():
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        flam.interp_decomp(_kahan(50, 0.7), 1e-6, Tmax=2, rrqr_iter=1)
    assert any("FLAM:id:maxIterCount" in str(x.message) for x in w)


# -------------------------------------------------------------- rskelf -----
def test_readme_rank_table():
    """README.md:99-106 with the demo's own matrix and settings."""
    g = _readme()
    p = g["params"]
    pr = problems.CurveDL(p["N"], "ie_circle", nproxy=p["p"])
    F = flam.rskelf(pr.afun, pr.x, p["occ"], p["rank_or_tol"], pr.pxyfun,
                    dict(Tmax=p["Tmax"], symm=p["symm"]))
    got = flam.rank_table(F)
    for (lvl, nblk, s, e), (rl, rn, rs, re) in zip(got, g["rows"]):
        assert (lvl, nblk) == (rl, rn)
        assert abs(e - re) <= max(2, 0.05 * re)          # BLAS/ordering sensitive
        assert abs(s - rs) <= max(2, 0.05 * rs)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(p["N"])
    x = flam.rskelf_sv(F, b)
    xe = pr.closed_form_inverse_apply(b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 10 * g["errors"]["rskelf_sv"]
    y = flam.rskelf_mv(F, b)
    assert np.linalg.norm(y - pr.matvec(b)) / np.linalg.norm(pr.matvec(b)) < 10 * g["errors"]["rskelf_mv"]


@pytest.mark.parametrize("order", ["sequential", "level"])
def test_circle_closed_form(order):
    """Benchmark circle matrix (alpha I + beta 11^T), Cholesky path."""
    N = 4096
    pr = problems.CurveDL(N, "circle")
    F = flam.rskelf(pr.afun, pr.x, 64, 1e-12, pr.pxyfun, dict(symm="p", order=order))
    rng = np.random.default_rng(0)
    b = rng.standard_normal(N)
    x = flam.rskelf_sv(F, b)
    xe = pr.closed_form_inverse_apply(b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-10
    assert np.allclose(pr.matvec(xe), b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_level_order_matches_sequential_ranks():
    """The snapshot-per-level order the device uses and the reference's
    sequential order differ only in which neighbour lists a box sees;
    per-level ranks agree to a few percent."""
    N = 8192
    pr = problems.CurveDL(N, "circle")
    tabs = {}
    for order in ("sequential", "level"):
        F = flam.rskelf(pr.afun, pr.x, 64, 1e-9, pr.pxyfun, dict(symm="p", order=order))
        tabs[order] = flam.rank_table(F)
    for a, b in zip(tabs["sequential"], tabs["level"]):
        assert a[0] == b[0] and a[1] == b[1]
        assert abs(a[3] - b[3]) <= max(4, 0.05 * a[3])


def test_ellipse_dense_lu():
    N = 1536
    pr = problems.CurveDL(N, "ellipse", a=2.0, b=1.0)
    A = pr.dense()
    F = flam.rskelf(pr.afun, pr.x, 64, 1e-10, pr.pxyfun, dict(symm="n"))
    rng = np.random.default_rng(1)
    B = rng.standard_normal((N, 3))
    X = flam.rskelf_sv(F, B)
    Xe = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xe) / np.linalg.norm(Xe) < 1e-7
    Y = flam.rskelf_mv(F, B)
    assert np.linalg.norm(Y - A @ B) / np.linalg.norm(A @ B) < 1e-7


def _factor(symm, cplx=False, N=1024):
    pr = problems.CurveDL(N, "circle")
    scale = (1.0 + 0.5j) if cplx else 1.0
    afun = (lambda I, J: pr.afun(I, J) * scale)

    def pxy(x, slf, nbr, l, ctr):
        K, nb = pr.pxyfun(x, slf, nbr, l, ctr)
        return K * scale, nb
    return flam.rskelf(afun, pr.x, 32, 1e-11, pxy, dict(symm=symm)), pr


@pytest.mark.parametrize("symm,cplx", [("n", False), ("p", False), ("h", False),
                                       ("n", True), ("s", True)])
def test_sweep_identities(symm, cplx):
    F, pr = _factor(symm, cplx)
    N = pr.N
    rng = np.random.default_rng(2)
    X = rng.standard_normal((N, 2)) + (1j * rng.standard_normal((N, 2)) if cplx else 0)
    Y = rng.standard_normal((N, 2)) + (1j * rng.standard_normal((N, 2)) if cplx else 0)
    for tr in ("n", "c", "t"):
        Z = flam.rskelf_sv(F, flam.rskelf_mv(F, X, tr), tr)
        assert np.linalg.norm(Z - X) / np.linalg.norm(X) < 1e-12
    # adjoint: <Y, F X> = <F^H Y, X>;  't' = conj('c' on conj)
    lhs = np.vdot(Y, flam.rskelf_mv(F, X, "n"))
    rhs = np.vdot(flam.rskelf_mv(F, Y, "c"), X)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    lhs = np.vdot(Y, flam.rskelf_sv(F, X, "n"))
    rhs = np.vdot(flam.rskelf_sv(F, Y, "c"), X)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    Zt = flam.rskelf_mv(F, X, "t")
    assert np.allclose(Zt, np.conj(flam.rskelf_mv(F, np.conj(X), "c")))
    # F approximates A
    scale = (1.0 + 0.5j) if cplx else 1.0
    AX = pr.matvec(X.real) * scale + (pr.matvec(X.imag) * scale * 1j if cplx else 0)
    assert np.linalg.norm(flam.rskelf_mv(F, X) - AX) / np.linalg.norm(AX) < 1e-9


def test_symm_modes_agree_on_spd():
    Fn, pr = _factor("n")
    Fp, _ = _factor("p")
    Fh, _ = _factor("h")
    b = np.random.default_rng(3).standard_normal(pr.N)
    xn, xp, xh = (flam.rskelf_sv(F, b) for F in (Fn, Fp, Fh))
    assert np.linalg.norm(xn - xp) < 1e-10 * np.linalg.norm(xn)
    assert np.linalg.norm(xn - xh) < 1e-10 * np.linalg.norm(xn)


def test_chol_fails_on_indefinite():
    """rskelf.m:235: symm 'p' on the demo's negative-definite circle matrix."""
    pr = problems.CurveDL(512, "ie_circle")
    with pytest.raises(np.linalg.LinAlgError):
        flam.rskelf(pr.afun, pr.x, 32, 1e-10, pr.pxyfun, dict(symm="p"))


def test_chk_errors():
    with pytest.raises(flam.FlamError) as e:
        flam.chksymm("q")
    assert e.value.ident == "FLAM:chksymm:invalidSymm"
    with pytest.raises(flam.FlamError) as e:
        flam.chktrans("x")
    assert e.value.ident == "FLAM:chktrans:invalidTrans"
    assert flam.chksymm("P") == "p" and flam.chktrans("C") == "c"


def test_partial_factorization_schur_system():
    """opts.stop (rskelf.m:128,172,279-308): sweeps with mode 1, a dense
    solve of the remaining skeleton system, then mode 2, solve A x = b."""
    N = 2048
    pr = problems.CurveDL(N, "circle")
    F = flam.rskelf(pr.afun, pr.x, 32, 1e-12, pr.pxyfun, dict(symm="n", stop=3))
    assert "Si" in F and len(F["Si"]) > 0
    b = np.random.default_rng(4).standard_normal(N)
    y = flam.rskelf_sv(F, b, mode=1)
    Si = F["Si"]
    S = pr.afun(Si, Si) + F["S"].toarray()
    y[Si] = np.linalg.solve(S, y[Si])
    x = flam.rskelf_sv(F, y, mode=2)
    xe = pr.closed_form_inverse_apply(b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-10


def test_snorm_matches_svd():
    rng = np.random.default_rng(6)
    A = rng.standard_normal((60, 60))
    s = flam.snorm(60, lambda v: A @ v, lambda v: A.T @ v, tol=1e-10, niter_max=2000)
    assert abs(s - np.linalg.norm(A, 2)) < 1e-3 * s


# ---------------------------------------------------------------- sphere --
def test_trisphere_counts_and_geometry():
    """geom/trisphere_subdiv.m: 20 * 4^k faces, 10 * 4^k + 2 vertices; the
    triangles tile an inscribed polyhedron (total area below 4 pi and
    growing with k, above 0.99 * 4 pi at k = 4), centroids inside the unit
    ball, normals outward."""
    prev = 0.0
    for k in range(5):
        V, F = problems.trisphere_subdiv(20 * 4 ** k)
        assert F.shape[1] == 20 * 4 ** k and V.shape[1] == 10 * 4 ** k + 2
        assert np.allclose(np.linalg.norm(V, axis=0), 1.0)
        C, Nrm, A = problems.tri3geom(V, F)
        assert prev < A.sum() < 4 * np.pi       # refining the inscribed polyhedron
        prev = A.sum()
        assert np.all(np.linalg.norm(C, axis=0) < 1.0)
        assert np.allclose(np.sum(Nrm * C, axis=0) / np.linalg.norm(C, axis=0), 1.0, atol=0.05)
    assert prev > 0.99 * 4 * np.pi
    V, _ = problems.trisphere_subdiv(512, "v")
    assert V.shape[1] == 642


def test_self_panel_closed_form_vs_duffy_quadrature():
    """The centroid self-panel integral of 1/(4 pi r) in closed form against
    brute-force Duffy quadrature (the construction of quad/quad_sqtri3.m),
    and against the equilateral triangle's textbook value
    sqrt(3) a ln(2 + sqrt(3)) / (4 pi) ... for side a, centroid apex."""
    V, F = problems.trisphere_subdiv(320)
    C, _, _ = problems.tri3geom(V, F)
    d1 = problems.self_panel_sl(V, F, C)
    d2 = problems.self_panel_sl_quadrature(V, F, C)
    assert np.max(np.abs(d1 - d2) / d1) < 1e-9
    # equilateral triangle of side a: each edge at distance h = a / (2 sqrt 3)
    # from the centroid, ends at -a/2, a/2: 3 h 2 asinh(sqrt 3) / (4 pi)
    a = 0.7
    Ve = np.array([[0, 0, 0], [a, 0, 0], [a / 2, a * np.sqrt(3) / 2, 0]]).T
    Fe = np.array([[0], [1], [2]])
    Ce = Ve.mean(axis=1, keepdims=True)
    h = a / (2 * np.sqrt(3))
    assert np.isclose(problems.self_panel_sl(Ve, Fe, Ce)[0], 6 * h * np.arcsinh(np.sqrt(3)) / (4 * np.pi),
                      rtol=1e-14)


def test_sphere_rskelf_vs_dense_solve():
    """3D single layer on the 1280-triangle sphere: the oracle's factor
    solves to the dense LU's answer within the ID tolerance, and a
    compressing configuration (tol 1e-4) stays within its bound."""
    pr = problems.SphereSL(1280)
    A = pr.dense()
    b = np.random.default_rng(0).standard_normal(pr.N)
    xe = np.linalg.solve(A, b)
    for occ, tol, bound in ((64, 1e-10, 1e-9), (32, 1e-4, 1e-2)):
        F = flam.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="n", order="level"))
        x = flam.rskelf_sv(F, b)
        assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < bound
        y = flam.rskelf_mv(F, xe)
        assert np.linalg.norm(y - b) / np.linalg.norm(b) < bound
    tab = flam.rank_table(F)
    assert tab[-1][3] == 0 and tab[0][2] == pr.N
    assert sum(r[2] - r[3] for r in tab) == pr.N


# ------------------------------------------------- logdet / cholmv / cholsv --
def _circle_logdet(N, kind):
    """log det of alpha I + beta 1 1' = (N - 1) log alpha + log(alpha + N beta)."""
    alpha, beta = (0.5 - 1.0 / N, 0.5 / N) if kind == "circle" else (-0.5, -0.5 / N)
    return (N - 1) * np.log(complex(alpha)) + np.log(complex(alpha + N * beta))


@pytest.mark.parametrize("symm,kind", [("p", "circle"), ("h", "circle"), ("n", "circle"),
                                       ("h", "ie_circle"), ("n", "ie_circle")])
def test_logdet_closed_form(symm, kind):
    """rskelf_logdet against det(alpha I + beta 1 1') in closed form; for the
    negative definite ie_circle matrix the imaginary part is N pi mod 2 pi."""
    N = 1024
    pr = problems.CurveDL(N, kind)
    F = flam.rskelf(pr.afun, pr.x, 64, 1e-12, pr.pxyfun, dict(symm=symm, order="level"))
    ld = flam.rskelf_logdet(F)
    ref = _circle_logdet(N, kind)
    assert abs(ld.real - ref.real) < 1e-9 * abs(ref.real)
    assert abs(ld.imag - math.fmod(ref.imag, 2 * math.pi) % (2 * math.pi)) < 1e-9
    # and the dense determinant of the ellipse ('n', permutation signs count)
    if symm == "n" and kind == "circle":
        pe = problems.CurveDL(512, "ellipse", b=0.5)
        Fe = flam.rskelf(pe.afun, pe.x, 32, 1e-12, pe.pxyfun, dict(symm="n", order="level"))
        sign, lad = np.linalg.slogdet(pe.dense())
        lde = flam.rskelf_logdet(Fe)
        assert abs(lde.real - lad) < 1e-9 * abs(lad)
        assert abs(np.exp(1j * lde.imag) - sign) < 1e-9


def test_cholmv_cholsv_identities():
    """C C' = F: cholmv(cholmv(x, 'c'), 'n') = rskelf_mv(x); cholsv inverts
    cholmv; 't' is the conjugate of 'c'; both refuse symm 'n'."""
    N = 1024
    pr = problems.CurveDL(N, "circle")
    F = flam.rskelf(pr.afun, pr.x, 64, 1e-12, pr.pxyfun, dict(symm="p", order="level"))
    x = np.random.default_rng(0).standard_normal((N, 2))
    y = flam.rskelf_cholmv(F, flam.rskelf_cholmv(F, x, "c"), "n")
    assert np.allclose(y, flam.rskelf_mv(F, x), atol=1e-13)
    assert np.allclose(flam.rskelf_cholsv(F, flam.rskelf_cholmv(F, x, "n"), "n"), x, atol=1e-12)
    assert np.allclose(flam.rskelf_cholsv(F, flam.rskelf_cholmv(F, x, "c"), "c"), x, atol=1e-12)
    assert np.allclose(flam.rskelf_cholmv(F, x, "t"), flam.rskelf_cholmv(F, x, "c"))
    # ||C' x||^2 = x' A x for the SPD matrix
    cx = flam.rskelf_cholmv(F, x[:, 0], "c")
    assert np.isclose(cx @ cx, x[:, 0] @ pr.matvec(x[:, 0]), rtol=1e-10)
    Fn = flam.rskelf(pr.afun, pr.x, 64, 1e-12, pr.pxyfun, dict(symm="n", order="level"))
    with pytest.raises(flam.FlamError):
        flam.rskelf_cholmv(Fn, x)


# ------------------------------------------------- matrix definitions -----
@pytest.mark.parametrize("kind,kw", [("ie_circle", {}), ("ellipse", dict(a=2.0, b=1.0)),
                                     ("ellipse", dict(a=1.0, b=0.3))])
def test_double_layer_gauss_identity(kind, kw):
    """Gauss's lemma: the double layer of a constant density is -1/2 on a
    smooth closed curve, so with the jump term -1/2 the interior Dirichlet
    matrix of ie_circle.m / ie_ellipse.m (trapezoidal weights, curvature
    diagonal -1/2 (1 + h kappa / 2 pi)) has every row sum -1, to spectral
    accuracy.  A flipped normal gives row sums near 0, a wrong curvature
    term misses by h kappa / 2 pi ~ 1e-3."""
    for N in (512, 2048):
        pr = problems.CurveDL(N, kind, **kw)
        assert np.abs(pr.dense().sum(axis=1) + 1.0).max() < 1e-12
    bad = problems.CurveDL(512, kind, **kw)
    bad.nu = -bad.nu
    assert np.abs(bad.dense().sum(axis=1) + 1.0).max() > 0.5
    if kind == "ellipse":                           # ie_circle.m:102 has no kappa term
        bad = problems.CurveDL(512, kind, **kw)
        bad.kappa = -bad.kappa
        assert np.abs(bad.dense().sum(axis=1) + 1.0).max() > 1e-4


@pytest.mark.parametrize("kind,kw,ratio,symm", [("ie_circle", {}, 1.0, "h"),
                                                ("ellipse", dict(a=2.0, b=1.0), 2.0, "n")])
def test_pde_known_answer_curves(kind, kw, ratio, symm):
    """ie_circle.m:65-81 / ie_ellipse.m:83-99: the field of exterior charges
    recovered inside from the solved double-layer density, through a dense
    solve and through the oracle's factor (README.md:111 reports 1.35e-12
    at N = 16384).  A sign error in the kernel or the jump term gives O(1)."""
    N = 2048
    pr = problems.CurveDL(N, kind, **kw)
    A = pr.dense()
    assert problems.laplace2d_pde_error(pr.x, pr.nu, pr.h, lambda b: np.linalg.solve(A, b),
                                        ratio) < 1e-13
    F = flam.rskelf(pr.afun, pr.x, 64, 1e-12, pr.pxyfun, dict(symm=symm))
    assert problems.laplace2d_pde_error(pr.x, pr.nu, pr.h, lambda b: flam.rskelf_sv(F, b),
                                        ratio) < 1e-11
    flipped = -A + np.diag(2 * np.diag(A))           # the exterior-jump sign
    assert problems.laplace2d_pde_error(pr.x, pr.nu, pr.h,
                                        lambda b: np.linalg.solve(flipped, b), ratio) > 1e-2


def test_sphere_single_layer_identities():
    """The single layer of a unit density on the unit sphere is 1 on the
    sphere (int dS_y / (4 pi |x - y|) = R), so the matrix's row sums tend to
    1 at first order in the mesh width (one-point quadrature, self-panel
    integral on the diagonal); the known-answer field of exterior charges
    (the ie_sphere.m:127-145 check for this kernel) converges at the same
    rate.  A missing or mis-scaled diagonal (~0.03 at 1280 triangles) or a
    wrong weight fails both."""
    errs, pde = [], []
    for nf in (1280, 5120):
        pr = problems.SphereSL(nf)
        A = pr.dense()
        errs.append(np.abs(A.sum(axis=1) - 1.0).max())
        pde.append(problems.sphere_pde_error(pr.x, pr.area, lambda b: np.linalg.solve(A, b)))
        if nf == 1280:
            Ad = A - np.diag(np.diag(A))
            assert np.abs(Ad.sum(axis=1) - 1.0).max() > 5 * errs[0]
            F = flam.rskelf(pr.afun, pr.x, 64, 1e-6, pr.pxyfun, dict(symm="n", order="level"))
            assert abs(problems.sphere_pde_error(pr.x, pr.area, lambda b: flam.rskelf_sv(F, b))
                       - pde[0]) < 1e-5
    assert errs[0] < 5e-3 and errs[1] < 2e-3 and errs[0] / errs[1] > 1.8
    assert pde[0] < 4e-3 and pde[1] < 2e-3 and pde[0] / pde[1] > 1.7


def test_symm_h_bunch_kaufman_on_indefinite():
    """rskelf.m:229's ldl pivots (Bunch-Kaufman): on an indefinite matrix
    whose first pivot is tiny against its column the oracle's 'h' solves
    to the dense answer, where an unpivoted L D L' would grow past 1e8."""
    N = 64
    pr = problems.CurveDL(N, "circle")
    d = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    d[0] = 1e-12 * 0.5 / N
    A = pr.dense()
    A[np.arange(N), np.arange(N)] = d
    afun = lambda I, J: A[np.ix_(np.asarray(I, dtype=np.int64), np.asarray(J, dtype=np.int64))]
    F = flam.rskelf(afun, pr.x, N, 1e-12, pr.pxyfun, dict(symm="h"))
    b = np.random.default_rng(5).standard_normal(N)
    xe = np.linalg.solve(A, b)
    assert np.linalg.norm(flam.rskelf_sv(F, b) - xe) <= 1e-10 * np.linalg.norm(xe)
    # unpivoted L D L' of the same matrix: L grows like 1 / A_00
    L = np.eye(N)
    L[1:, 0] = A[1:, 0] / A[0, 0]
    assert np.abs(L).max() > 1e8


# ------------------------------------------------- Helmholtz CFIE (complex) --
def test_cfie_circle_eigenvalues():
    """On the unit circle the combined-field matrix is circulant and complex
    symmetric, and e^{imt} is an eigenvector with the continuous
    operator's eigenvalue (iπk/2) J_m'(k) H_m(k) + (ηπ/2) J_m(k) H_m(k) up to
    the discretization error, which falls at third order in h (corrected
    trapezoidal rule).  This pins the kernel (scipy's hankel1), the jump
    term, the normal's direction and the sign of -i eta S."""
    errs = []
    for N in (256, 1024):
        pr = problems.CurveCFIE(N, "circle", k=10.0)
        A = pr.dense()
        assert np.abs(A - A.T).max() < 1e-13
        t = np.arange(1, N + 1) * 2 * np.pi / N
        e = 0.0
        for m in range(-25, 26):
            v = np.exp(1j * m * t)
            lam = pr.circle_eigenvalue(m)
            e = max(e, np.linalg.norm(A @ v - lam * v) / (np.linalg.norm(v) * abs(lam)))
        errs.append(e)
    assert errs[1] < 5e-5 and errs[0] / errs[1] > 30
    bad = problems.CurveCFIE(1024, "circle", k=10.0)
    bad.nu = -bad.nu
    v = np.exp(1j * 3 * np.arange(1, 1025) * 2 * np.pi / 1024)
    lam = pr.circle_eigenvalue(3)
    assert np.linalg.norm(bad.dense() @ v - lam * v) / (np.linalg.norm(v) * abs(lam)) > 0.1


def test_cfie_star_pde_known_answer():
    """Exterior known answer on the star curve r = 1 + 0.3 cos 5t: the field
    of a point source inside, recovered outside from the solved density,
    with third-order convergence; through a dense solve and the oracle's
    factor (tol 1e-10)."""
    errs = []
    for N in (512, 2048):
        pr = problems.CurveCFIE(N, "star", k=10.0)
        A = pr.dense()
        errs.append(problems.helmholtz_pde_error(pr, lambda b: np.linalg.solve(A, b)))
    assert errs[1] < 1e-6 and errs[0] / errs[1] > 30
    F = flam.rskelf(pr.afun, pr.x, 64, 1e-10, pr.pxyfun, dict(symm="n", order="level"))
    assert abs(problems.helmholtz_pde_error(pr, lambda b: flam.rskelf_sv(F, b)) - errs[1]) < 1e-8


@pytest.mark.parametrize("kind,symm", [("star", "n"), ("circle", "s")])
def test_cfie_rskelf_vs_dense(kind, symm):
    """Complex factorization ('n': LU after two-sided compression with
    conjugate-transposed incoming rows; 's': one-sided compression and
    T.' updates, rskelf.m:220-221, for the complex symmetric circle):
    solves for trans 'n', 'c', 't' and the apply against the dense matrix."""
    N, tol = 2048, 1e-9
    pr = problems.CurveCFIE(N, kind, k=20.0, symm=symm)
    F = flam.rskelf(pr.afun, pr.x, 64, tol, pr.pxyfun, dict(symm=symm, order="level"))
    A = pr.dense()
    rng = np.random.default_rng(3)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for trans, M in (("n", A), ("c", A.conj().T), ("t", A.T)):
        xe = np.linalg.solve(M, b)
        assert np.linalg.norm(flam.rskelf_sv(F, b, trans) - xe) < 100 * tol * np.linalg.norm(xe)
        ye = M @ b
        assert np.linalg.norm(flam.rskelf_mv(F, b, trans) - ye) < 100 * tol * np.linalg.norm(ye)


def test_fgmres_restatement():
    """The CPU FGMRES (the benchmark's CPU leg) with a loose factor as the
    preconditioner reaches the requested true residual on every column and
    agrees with the closed-form solution."""
    N = 4096
    pr = problems.CurveDL(N, "circle")
    F = flam.rskelf(pr.afun, pr.x, 64, 1e-4, pr.pxyfun, dict(symm="p", order="level"))
    B = np.random.default_rng(2).standard_normal((N, 3))
    X, iters, rel = flam.fgmres(pr.matvec, B, lambda V: flam.rskelf_sv(F, V), tol=1e-12)
    assert np.all(rel <= 1e-12) and iters <= 8
    Xe = pr.closed_form_inverse_apply(B)
    assert np.linalg.norm(X - Xe) <= 1e-10 * np.linalg.norm(Xe)


def test_mlp_restatement_gradient_by_finite_differences():
    """The rank model's CPU restatement: back-propagated gradient against
    central differences of the loss (the device gradient is checked against
    this restatement on the GPU)."""
    from oracle import mlp
    rng = np.random.default_rng(0)
    p = 0.3 * rng.standard_normal(mlp.NPARAMS)
    X = rng.standard_normal((40, mlp.F))
    t = rng.standard_normal(40)
    loss, g = mlp.loss_grad(p, X, t)
    h = 1e-6
    idx = rng.choice(mlp.NPARAMS, 60, replace=False).tolist() + [mlp.NPARAMS - 1, 0]
    for i in idx:
        e = np.zeros(mlp.NPARAMS)
        e[i] = h
        fd = (mlp.loss_grad(p + e, X, t)[0] - mlp.loss_grad(p - e, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i]))
