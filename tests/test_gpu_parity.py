"""Parity of the HIP rskelf path with the CPU oracle (oracle/flam.py).

Every call goes through the C ABI (flam_amd.rskelf -> librskx.so).  The
oracle runs in 'level' order: all boxes of a level see their neighbours'
uncompressed point lists, the convention of a level-parallel device
factorization (the reference's sequential order differs by a few skeletons
per level; tests/test_oracle.py pins that order against README.md).

Tolerances: ranks are data-dependent decisions of a pivoted QR; on
symmetric geometries ties can break differently, so per-level skeleton
totals must agree within 2% and at least 90% of boxes must have the same
rank.  Solves must agree with the closed form / dense LU / oracle to 1e-8
in FP64 (the factor tolerance is 1e-10 .. 1e-12).
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import flam as ofl  # noqa: E402
from oracle import problems as opr  # noqa: E402


def _dev():
    import flam_amd
    from flam_amd import geometry
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return flam_amd, geometry


def _cmp_tables(got, ref, frac=0.02):
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert a[:3] == b[:3] or (a[0], a[1]) == (b[0], b[1]), (a, b)
        assert abs(a[3] - b[3]) <= max(2, frac * b[3]), (a, b)


def _box_ranks(Fd, Fo):
    n = len(Fo["factors"])
    ko = np.array([len(f["sk"]) for f in Fo["factors"]])
    kd = np.array([Fd.export(i)["sk"].shape[0] for i in range(Fd.info()["nfactor"])])
    return kd, ko


@pytest.mark.parametrize("symm", ["p", "n"])
def test_circle_fp64_vs_oracle_and_closed_form(symm):
    flam_amd, geometry = _dev()
    N, occ, tol = 4096, 64, 1e-12
    print(f"circle N={N} occ={occ} tol={tol} symm={symm}", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm=symm))
    pr = opr.CurveDL(N, "circle")
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm=symm, order="level"))
    _cmp_tables(F.level_table(), ofl.rank_table(Fo))
    kd, ko = _box_ranks(F, Fo)
    assert len(kd) == len(ko)
    assert np.mean(kd == ko) >= 0.9
    rng = np.random.default_rng(0)
    B = rng.standard_normal((N, 3))
    X = flam_amd.rskelf_sv(F, torch.from_numpy(B).cuda()).cpu().numpy()
    Xe = pr.closed_form_inverse_apply(B)
    assert np.linalg.norm(X - Xe) / np.linalg.norm(Xe) < 1e-10
    Xo = ofl.rskelf_sv(Fo, B)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10
    Y = flam_amd.rskelf_mv(F, torch.from_numpy(B).cuda()).cpu().numpy()
    AB = pr.matvec(B)
    assert np.linalg.norm(Y - AB) / np.linalg.norm(AB) < 1e-10


def test_ellipse_nonsymmetric_vs_dense_lu():
    flam_amd, geometry = _dev()
    N, occ, tol = 2048, 64, 1e-10
    print(f"ellipse N={N} a=2 b=1 occ={occ} tol={tol} symm=n", flush=True)
    g = geometry.ellipse(N, a=2.0, b=1.0)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="n"))
    pr = opr.CurveDL(N, "ellipse", a=2.0, b=1.0)
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="n", order="level"))
    _cmp_tables(F.level_table(), ofl.rank_table(Fo))
    A = pr.dense()
    rng = np.random.default_rng(1)
    B = rng.standard_normal((N, 2))
    X = flam_amd.rskelf_sv(F, torch.from_numpy(B).cuda()).cpu().numpy()
    Xe = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xe) / np.linalg.norm(Xe) < 1e-7
    Xo = ofl.rskelf_sv(Fo, B)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-7
    for tr in ("n", "c", "t"):
        Y = flam_amd.rskelf_mv(F, torch.from_numpy(B).cuda(), tr).cpu().numpy()
        Yo = ofl.rskelf_mv(Fo, B, tr)
        assert np.linalg.norm(Y - Yo) / np.linalg.norm(Yo) < 1e-7, tr
        Z = flam_amd.rskelf_sv(F, torch.from_numpy(Y).cuda(), tr).cpu().numpy()
        assert np.linalg.norm(Z - B) / np.linalg.norm(B) < 1e-11, tr
    # the device's F equals the oracle's F as an operator on the sampled
    # rows (ie_ellipse.m:48-81 checks 128 random rows against A)
    Ad = flam_amd.kernel_matvec(g.kernel, g.x, torch.from_numpy(B).cuda()).cpu().numpy()
    assert np.linalg.norm(Ad - A @ B) / np.linalg.norm(A @ B) < 1e-13


def test_factor_export_matches_oracle_structure():
    """F.factors(i): same rank and skeleton set per box where ranks agree;
    L U = K(rd(p), rd) holds up to rounding for the exported blocks."""
    flam_amd, geometry = _dev()
    N, occ, tol = 2048, 64, 1e-10
    g = geometry.ellipse(N, a=1.5, b=1.0)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="n"))
    pr = opr.CurveDL(N, "ellipse", a=1.5, b=1.0)
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="n", order="level"))
    nf = F.info()["nfactor"]
    assert nf == len(Fo["factors"])
    same = 0
    for i in range(nf):
        fd, fo = F.export(i), Fo["factors"][i]
        assert sorted(np.concatenate([fd["sk"], fd["rd"]])) == sorted(np.concatenate([fo["sk"], fo["rd"]]))
        if len(fd["sk"]) == len(fo["sk"]) and set(fd["sk"]) == set(fo["sk"]):
            same += 1
        assert np.abs(fd["T"]).max(initial=0) <= 2 + 1e-8          # Tmax = 2
        L, U, p = fd["L"], fd["U"], fd["p"]
        assert np.allclose(np.tril(L), L) and np.allclose(np.triu(U), U)
        assert np.all(np.sort(p) == np.arange(len(p)))
    assert same >= 0.85 * nf


def test_mixed_precision_refined_to_1e10():
    """FP32 factor at tol 1e-6 as the preconditioner of FP64 flexible
    GMRES reaches ||A x - b|| / ||b|| <= 1e-10."""
    flam_amd, geometry = _dev()
    from flam_amd.refine import fgmres
    N, occ, tol = 16384, 128, 1e-6
    print(f"mixed circle N={N} occ={occ} tol={tol} fp32 + fgmres", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="p", precision="fp32"))
    rng = np.random.default_rng(0)
    B = torch.from_numpy(rng.standard_normal((N, 4))).cuda()
    X, info = fgmres(g.circulant_matvec, B, lambda V: flam_amd.rskelf_sv(F, V), tol=1e-10)
    R = B - g.circulant_matvec(X)
    rel = (torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(B, dim=0)).max().item()
    print("fgmres", info, "relres", rel, flush=True)
    assert rel <= 1e-10
    pr = opr.CurveDL(N, "circle")
    Xe = pr.closed_form_inverse_apply(B.cpu().numpy())
    assert np.linalg.norm(X.cpu().numpy() - Xe) / np.linalg.norm(Xe) < 1e-9


def test_fp32_factor_ranks_match_fp64_oracle():
    """FP64 ID decisions on FP32-stored blocks: ranks per level within 2%
    of the FP64 oracle at tol 1e-6."""
    flam_amd, geometry = _dev()
    N, occ, tol = 8192, 128, 1e-6
    g = geometry.ellipse(N, a=1.0, b=0.6)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="n", precision="fp32"))
    pr = opr.CurveDL(N, "ellipse", a=1.0, b=0.6)
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="n", order="level"))
    _cmp_tables(F.level_table(), ofl.rank_table(Fo), frac=0.03)


@pytest.mark.parametrize("N,occ", [(1, 4), (5, 8), (37, 4), (300, 300)])
def test_edge_sizes(N, occ):
    """Single box (root only: K empty, full dense factor), tiny trees and
    a box at exactly occ points."""
    flam_amd, geometry = _dev()
    g = geometry.ie_circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, occ, 1e-12, g.proxy, dict(symm="n"))
    pr = opr.CurveDL(N, "ie_circle")
    b = np.random.default_rng(2).standard_normal(N)
    x = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda()).cpu().numpy()
    xe = np.linalg.solve(pr.dense(), b)
    assert np.linalg.norm(x - xe) <= 1e-9 * np.linalg.norm(xe)


def test_refinement_off_and_iteration_cap():
    """Tmax = Inf skips refinement; rrqr_iter = 0 too; both still solve."""
    flam_amd, geometry = _dev()
    N = 2048
    g = geometry.ellipse(N, a=2.0, b=1.0)
    pr = opr.CurveDL(N, "ellipse", a=2.0, b=1.0)
    b = np.random.default_rng(3).standard_normal(N)
    xe = np.linalg.solve(pr.dense(), b)
    for o in (dict(Tmax=math.inf), dict(rrqr_iter=0)):
        F = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm="n", **o))
        x = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda()).cpu().numpy()
        assert np.linalg.norm(x - xe) <= 1e-7 * np.linalg.norm(xe)


def test_chol_failure_reported():
    """symm 'p' on the demo's negative-definite matrix (rskelf.m:235)."""
    flam_amd, geometry = _dev()
    g = geometry.ie_circle(1024)
    with pytest.raises(flam_amd.FlamError) as e:
        flam_amd.rskelf(g.kernel, g.x, 32, 1e-10, g.proxy, dict(symm="p"))
    assert e.value.ident == "FLAM:rskx:chol"


def test_readme_configuration_rank_table():
    """The README demo exactly as README.md:95-111 ran it (ie_circle,
    N=16384, occ=64, tol=1e-12, symm 'h', the ie_circle.m:31 default) on
    the device: block counts exact, ranks within 6% of README.md:99-106,
    and the demo's PDE known-answer error at the README's level
    (README.md:111: 1.35e-12)."""
    import json
    import os
    flam_amd, geometry = _dev()
    with open(os.path.join(os.path.dirname(__file__), "golden", "readme_ie_circle.json")) as f:
        gold = json.load(f)
    g = geometry.ie_circle(16384)
    F = flam_amd.rskelf(g.kernel, g.x, 64, 1e-12, g.proxy, dict(symm="h"))
    for (lvl, nblk, s, e), (rl, rn, rs, re) in zip(F.level_table(), gold["rows"]):
        assert (lvl, nblk) == (rl, rn)
        assert abs(e - re) <= max(2, 0.06 * re)
    x, nu, w = (t.cpu().numpy() for t in (g.x, g.nu, g.w))
    err = opr.laplace2d_pde_error(
        x, nu, w, lambda B: flam_amd.rskelf_sv(F, torch.from_numpy(B).cuda()).cpu().numpy())
    print(f"readme configuration: pde error {err:.3e}", flush=True)
    assert err < 1e-11


def test_sweeps_leave_input_untouched():
    """rskelf_sv / rskelf_mv return new arrays (MATLAB value semantics);
    a 1-D CUDA input must not be overwritten by the in-place device sweep."""
    flam_amd, geometry = _dev()
    g = geometry.circle(1024)
    F = flam_amd.rskelf(g.kernel, g.x, 32, 1e-10, g.proxy, dict(symm="p"))
    for shape in ((1024,), (1024, 1), (1024, 3)):
        b = torch.randn(*shape, dtype=torch.float64, device="cuda")
        keep = b.clone()
        x = flam_amd.rskelf_sv(F, b)
        y = flam_amd.rskelf_mv(F, b)
        assert torch.equal(b, keep)
        assert x.shape == b.shape and y.shape == b.shape
        assert not torch.equal(x, b)


@pytest.mark.parametrize("shape,symm", [("circle", "p"), ("ellipse", "n")])
def test_partial_factorization(shape, symm):
    """opts.stop (rskelf.m:128,172,279-308) and rskelf_partial_info/_sv/_mv:
    as many points remain as in the oracle (within the rank tolerance of
    the tables above); when the remaining sets coincide the skeleton system
    S agrees to the accuracy the interpolation matrices allow (the skeleton
    blocks R11 have condition up to ~1/tol, so T, and through it S, is
    reproduced only to ~1e-16/tol); and the partial sweeps around a dense
    solve (product) of A(sk,sk) + S reproduce A^{-1} b (A b) to ~tol."""
    flam_amd, geometry = _dev()
    N, occ, tol, stop = 2048, 32, 1e-10, 3
    print(f"partial {shape} {symm} N={N} stop={stop}", flush=True)
    if shape == "circle":
        g, pr = geometry.circle(N), opr.CurveDL(N, "circle")
    else:
        g, pr = geometry.ellipse(N, a=1.0, b=0.6), opr.CurveDL(N, "ellipse", a=1.0, b=0.6)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm=symm, stop=stop))
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm=symm, stop=stop, order="level"))
    assert F.info()["nlvl_done"] == stop
    sk, S = flam_amd.rskelf_partial_info(F)
    Sd = S.to_dense().numpy()
    assert abs(sk.numel() - len(Fo["Si"])) <= max(2, 0.02 * len(Fo["Si"]))
    assert np.all(np.diff(sk.numpy()) > 0)
    if np.array_equal(sk.numpy(), np.asarray(Fo["Si"])):
        So = Fo["S"].toarray()
        assert np.abs(Sd - So).max() <= 1e-2 * np.abs(So).max()
    skn = sk.numpy()
    Ask = pr.afun(skn, skn) + Sd
    b = np.random.default_rng(5).standard_normal((N, 2))
    bt = torch.from_numpy(b).cuda()
    Ad = pr.dense()
    for trans in ("n", "c"):
        Aop = Ad if trans == "n" else Ad.T
        Askop = Ask if trans == "n" else Ask.T
        x = flam_amd.rskelf_partial_sv(
            F, bt, lambda Y, t: torch.from_numpy(np.linalg.solve(Askop, Y.cpu().numpy())).cuda(),
            trans).cpu().numpy()
        xe = np.linalg.solve(Aop, b)
        assert np.linalg.norm(x - xe) <= 1e-7 * np.linalg.norm(xe), trans
        y = flam_amd.rskelf_partial_mv(
            F, bt, lambda Y, t: torch.from_numpy(Askop @ Y.cpu().numpy()).cuda(), trans).cpu().numpy()
        ye = Aop @ b
        assert np.linalg.norm(y - ye) <= 1e-7 * np.linalg.norm(ye), trans


def test_partial_factorization_stop_zero():
    """stop = 0: nothing is factored; every point remains and S is empty."""
    flam_amd, geometry = _dev()
    N = 600
    g = geometry.ie_circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, 32, 1e-10, g.proxy, dict(symm="n", stop=0))
    sk, S = flam_amd.rskelf_partial_info(F)
    assert F.info()["nlvl_done"] == 0 and F.info()["nfactor"] == 0
    assert np.array_equal(sk.numpy(), np.arange(N)) and S._nnz() == 0


# ------------------------------------------------------------ symm 'h' --
@pytest.mark.parametrize("kind", ["circle", "ie_circle"])
def test_symm_h_ldl(kind):
    """symm 'h' (rskelf.m:228-233, sv/rskelf_sv_h.m, mv/rskelf_mv_h.m): the
    same ID as 'p' (identical rank table to the oracle's 'h'), L D L'
    with Bunch-Kaufman pivots on the SPD benchmark circle and on the demo's negative
    definite ie_circle matrix (where Cholesky fails); solve against the
    closed form, apply against the exact circulant product, exported L unit
    lower and U = D diagonal, and 't' / 'c' equal to 'n'."""
    flam_amd, geometry = _dev()
    N, occ, tol = 4096, 64, 1e-12
    print(f"symm h: {kind} N={N} occ={occ} tol={tol}", flush=True)
    g = geometry.circle(N) if kind == "circle" else geometry.ie_circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="h"))
    pr = opr.CurveDL(N, kind)
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="h", order="level"))
    _cmp_tables(F.level_table(), ofl.rank_table(Fo))
    b = np.random.default_rng(5).standard_normal(N)
    bt = torch.from_numpy(b).cuda()
    x = flam_amd.rskelf_sv(F, bt).cpu().numpy()
    xe = pr.closed_form_inverse_apply(b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-10
    y = flam_amd.rskelf_mv(F, torch.from_numpy(xe).cuda()).cpu().numpy()
    assert np.linalg.norm(y - b) / np.linalg.norm(b) < 1e-10
    for t in ("t", "c"):
        assert torch.allclose(flam_amd.rskelf_sv(F, bt, t), flam_amd.rskelf_sv(F, bt), rtol=0,
                              atol=1e-14 * np.abs(x).max())
    f = F.export(0)
    assert np.allclose(np.diag(f["L"]), 1.0) and np.allclose(np.triu(f["L"], 1), 0.0)
    assert np.allclose(f["U"], np.diag(np.diag(f["U"])))
    d = np.diag(f["U"])
    assert (d > 0).all() if kind == "circle" else (d < 0).all()
    if kind == "ie_circle":
        with pytest.raises(flam_amd.FlamError) as e:
            flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="p"))
        assert e.value.ident == "FLAM:rskx:chol"
    # boxes above the register elimination's size (the LDS / global kernel)
    F2 = flam_amd.rskelf(g.kernel, g.x, 300, tol, g.proxy, dict(symm="h"))
    x2 = flam_amd.rskelf_sv(F2, bt).cpu().numpy()
    assert np.linalg.norm(x2 - xe) / np.linalg.norm(xe) < 1e-10
    y2 = flam_amd.rskelf_mv(F2, torch.from_numpy(xe).cuda()).cpu().numpy()
    assert np.linalg.norm(y2 - b) / np.linalg.norm(b) < 1e-10


def test_symm_h_fp32_refined():
    flam_amd, geometry = _dev()
    from flam_amd.refine import fgmres
    N = 1 << 16
    print(f"symm h fp32: circle N={N} occ=256 tol=1e-6 + fgmres", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, dict(symm="h", precision="fp32"))
    B = torch.from_numpy(np.random.default_rng(6).standard_normal((N, 3))).cuda()
    X, info = fgmres(g.circulant_matvec, B, lambda V: flam_amd.rskelf_sv(F, V), tol=1e-10)
    R = B - g.circulant_matvec(X)
    rel = (torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(B, dim=0)).max().item()
    assert rel <= 1e-10 and info["iters"] <= 8


# ------------------------------------------------- logdet / cholmv / cholsv --
@pytest.mark.parametrize("symm,kind,prec", [("p", "circle", "fp64"), ("h", "circle", "fp64"),
                                            ("n", "circle", "fp64"), ("h", "ie_circle", "fp64"),
                                            ("n", "ie_circle", "fp64"), ("p", "circle", "fp32")])
def test_logdet_vs_closed_form_and_oracle(symm, kind, prec):
    """rskelf_logdet (rskelf_logdet.m) against log det(alpha I + beta 1 1')
    and against the oracle's factor; FP32 factors to FP32 accuracy."""
    flam_amd, geometry = _dev()
    N, occ, tol = 4096, 64, 1e-12
    print(f"logdet {symm} {kind} {prec} N={N}", flush=True)
    g = geometry.circle(N) if kind == "circle" else geometry.ie_circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm=symm, precision=prec))
    ld = flam_amd.rskelf_logdet(F)
    alpha, beta = (0.5 - 1.0 / N, 0.5 / N) if kind == "circle" else (-0.5, -0.5 / N)
    ref = (N - 1) * np.log(complex(alpha)) + np.log(complex(alpha + N * beta))
    rtol = 1e-10 if prec == "fp64" else 1e-4
    assert abs(ld.real - ref.real) <= rtol * abs(ref.real), (ld, ref)
    assert abs(ld.imag - (ref.imag % (2 * math.pi))) < 1e-9
    if prec == "fp64":
        pr = opr.CurveDL(N, kind)
        Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm=symm, order="level"))
        ldo = ofl.rskelf_logdet(Fo)
        assert abs(ld.real - ldo.real) <= 1e-9 * abs(ref.real)
        assert abs(np.exp(1j * ld.imag) - np.exp(1j * ldo.imag)) < 1e-8     # phases mod 2 pi


def test_cholmv_cholsv():
    """C C' = F for symm 'p' (rskelf_cholmv.m, rskelf_cholsv.m): C (C' x) =
    rskelf_mv(x), cholsv inverts cholmv for 'n' and 'c', the halves agree
    with the oracle's, ||C' x||^2 = x' A x, and symm 'n' is refused."""
    flam_amd, geometry = _dev()
    N = 4096
    print(f"cholmv/cholsv circle N={N}", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, 64, 1e-12, g.proxy, dict(symm="p"))
    x = torch.from_numpy(np.random.default_rng(7).standard_normal((N, 3))).cuda()
    y = flam_amd.rskelf_cholmv(F, flam_amd.rskelf_cholmv(F, x, "c"), "n")
    assert torch.allclose(y, flam_amd.rskelf_mv(F, x), rtol=0, atol=1e-13)
    for t in ("n", "c"):
        z = flam_amd.rskelf_cholsv(F, flam_amd.rskelf_cholmv(F, x, t), t)
        assert torch.allclose(z, x, rtol=0, atol=1e-11)
    assert torch.equal(flam_amd.rskelf_cholmv(F, x, "t"), flam_amd.rskelf_cholmv(F, x, "c"))
    cx = flam_amd.rskelf_cholmv(F, x[:, 0], "c")
    ax = g.circulant_matvec(x[:, 0])
    assert abs((cx @ cx).item() - (x[:, 0] @ ax).item()) <= 1e-10 * (x[:, 0] @ ax).item()
    # C depends on the ID's skeleton choices (ties may break differently from
    # the oracle's), so compare what does not: ||C' x||^2 and ||C \ x||^2
    # against x' F x and x' F^{-1} x of the oracle's factor
    pr = opr.CurveDL(N, "circle")
    Fo = ofl.rskelf(pr.afun, pr.x, 64, 1e-12, pr.pxyfun, dict(symm="p", order="level"))
    xc = x[:, 0].cpu().numpy()
    cs = flam_amd.rskelf_cholsv(F, x[:, 0], "n")
    assert abs((cx @ cx).item() - xc @ ofl.rskelf_mv(Fo, xc)) <= 1e-10 * (cx @ cx).item()
    assert abs((cs @ cs).item() - xc @ ofl.rskelf_sv(Fo, xc)) <= 1e-10 * (cs @ cs).item()
    Fn = flam_amd.rskelf(g.kernel, g.x, 64, 1e-12, g.proxy, dict(symm="n"))
    with pytest.raises(flam_amd.FlamError) as e:
        flam_amd.rskelf_cholmv(Fn, x)
    assert e.value.ident == "FLAM:rskelf_cholmv:invalidSymm"


# ------------------------------------------------------------ symm 's' --
def test_symm_s_real_symmetric():
    """symm 's' on the real symmetric benchmark circle (rskelf.m:210,220-227,
    sv/rskelf_sv_sn.m, sv/rskelf_sv_sc.m): one-sided compression (the
    oracle's 's' rank table), LU with partial pivoting; solve against the
    closed form, apply against the exact circulant product, 't' / 'c'
    solves of the symmetric matrix, exported factors with a permutation,
    and boxes beyond the register kernel."""
    flam_amd, geometry = _dev()
    N, occ, tol = 4096, 64, 1e-12
    print(f"symm s: circle N={N} occ={occ} tol={tol}", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="s"))
    pr = opr.CurveDL(N, "circle")
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="s", order="level"))
    _cmp_tables(F.level_table(), ofl.rank_table(Fo))
    assert F.info()["symm"] == "s"
    b = np.random.default_rng(11).standard_normal(N)
    bt = torch.from_numpy(b).cuda()
    xe = pr.closed_form_inverse_apply(b)
    for t in ("n", "t", "c"):
        x = flam_amd.rskelf_sv(F, bt, t).cpu().numpy()
        assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-10, t
    y = flam_amd.rskelf_mv(F, torch.from_numpy(xe).cuda()).cpu().numpy()
    assert np.linalg.norm(y - b) / np.linalg.norm(b) < 1e-10
    f = F.export(0)
    assert f["p"] is not None and sorted(f["p"].tolist()) == list(range(len(f["rd"])))
    assert np.allclose(np.diag(f["L"]), 1.0) and np.allclose(np.triu(f["L"], 1), 0.0)
    F2 = flam_amd.rskelf(g.kernel, g.x, 300, tol, g.proxy, dict(symm="s"))
    x2 = flam_amd.rskelf_sv(F2, bt).cpu().numpy()
    assert np.linalg.norm(x2 - xe) / np.linalg.norm(xe) < 1e-10


# ------------------------------------------------------------- streams --
def test_work_on_a_non_default_torch_stream():
    """The library runs on torch's current stream (its sweeps only queue
    their work): a factorization and solves issued inside a side stream
    give the default stream's results bit for bit, with torch work on the
    same stream reading them without a host synchronization, and switching
    back drains the side stream first."""
    flam_amd, geometry = _dev()
    N = 4096
    print(f"streams: circle N={N} occ=64 tol=1e-10", flush=True)
    g = geometry.circle(N)
    B = torch.from_numpy(np.random.default_rng(12).standard_normal((N, 3))).cuda()
    F0 = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm="p"))
    X0 = flam_amd.rskelf_sv(F0, B)
    Y0 = flam_amd.rskelf_mv(F0, X0)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        F1 = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm="p"))
        X1 = flam_amd.rskelf_sv(F1, B)
        Y1 = flam_amd.rskelf_mv(F1, X1 * 1.0)       # torch work between two library calls
        D1 = (Y1 - B).abs().max()
    torch.cuda.current_stream().wait_stream(s)
    X2 = flam_amd.rskelf_sv(F1, B)                   # back on the default stream
    assert torch.equal(X1, X0) and torch.equal(Y1, Y0) and torch.equal(X2, X0)
    assert D1.item() <= 1e-8 * B.abs().max().item()


def test_release_cache_between_factorizations():
    """rskx_release_cache returns the workspaces and cached blocks; a live
    factor keeps working and a new factorization allocates afresh."""
    flam_amd, geometry = _dev()
    N = 4096
    print(f"release cache: circle N={N}", flush=True)
    g = geometry.circle(N)
    F1 = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm="p"))
    F2 = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm="p"))
    b = torch.from_numpy(np.random.default_rng(0).standard_normal(N)).cuda()
    x1 = flam_amd.rskelf_sv(F1, b)
    del F2
    flam_amd.release_cache()
    x1b = flam_amd.rskelf_sv(F1, b)
    assert torch.equal(x1, x1b)
    F3 = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm="p"))
    assert torch.equal(flam_amd.rskelf_sv(F3, b), x1)


def _indefinite_circle(geometry, N):
    """The benchmark circle's constant off-diagonal b = 1/(2N) with a chosen
    diagonal: +-1 alternating and A_00 = 1e-12 b, an indefinite matrix whose
    first pivot is tiny against its column (Bunch-Kaufman interchanges it
    with row 1 and takes a 1 x 1 pivot there)."""
    g = geometry.circle(N)
    d = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    d[0] = 1e-12 * 0.5 / N
    g.diag = torch.from_numpy(d).cuda()
    A = opr.CurveDL(N, "circle").dense()
    A[np.arange(N), np.arange(N)] = d
    return g, A


def _two_by_two_matrix(N):
    """The benchmark circle's off-diagonal b = 1/(2N) with diagonal +-1,
    except A_00 = A_11 = 1e-12 b and A_01 = A_10 = 1: column 0's largest
    entry is A_10 = 1 against a negligible diagonal and so is row 1's, so
    Bunch-Kaufman takes the 2 x 2 pivot {0, 1} (dsytf2's rule)."""
    A = opr.CurveDL(N, "circle").dense()
    d = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    b = 0.5 / N
    d[0] = d[1] = 1e-12 * b
    A[np.arange(N), np.arange(N)] = d
    A[0, 1] = A[1, 0] = 1.0
    return A


@pytest.mark.parametrize("N", [64, 300])
def test_symm_h_bunch_kaufman(N):
    """symm 'h' with rskelf.m:229's Bunch-Kaufman pivots, on indefinite
    matrices whose unpivoted L D L' would break down (one box, so K(rd,rd)
    is the whole matrix; N = 64 eliminates in shared memory, 300 in global
    scratch).  The tiny-pivot circle (a 1 x 1 pivot after an interchange)
    and a matrix needing a 2 x 2 pivot (given to the library as a host
    callback): solve and apply to the dense answers, L unit lower and
    bounded by 1 / (1 - alpha), L D L' = A(rd, rd) in the exported order,
    D block diagonal with the 2 x 2 block, logdet equal to the dense
    slogdet, and the same solve as the oracle's 'h' (SciPy's
    Bunch-Kaufman)."""
    flam_amd, geometry = _dev()
    print(f"symm h bunch-kaufman: N={N} (one box)", flush=True)
    g, A1 = _indefinite_circle(geometry, N)
    A2 = _two_by_two_matrix(N)
    pr = opr.CurveDL(N, "circle")
    afun2 = lambda I, J: A2[np.ix_(np.asarray(I, dtype=np.int64), np.asarray(J, dtype=np.int64))]
    alpha = (1 + np.sqrt(17)) / 8
    b = np.random.default_rng(5).standard_normal(N)
    for A, F, blocks in (
            (A1, flam_amd.rskelf(g.kernel, g.x, N, 1e-12, g.proxy, dict(symm="h")), False),
            (A2, flam_amd.rskelf(afun2, pr.x, N, 1e-12, pr.pxyfun, dict(symm="h")), True)):
        x = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda()).cpu().numpy()
        xe = np.linalg.solve(A, b)
        assert np.linalg.norm(x - xe) <= 1e-10 * np.linalg.norm(xe)
        y = flam_amd.rskelf_mv(F, torch.from_numpy(xe).cuda()).cpu().numpy()
        assert np.linalg.norm(y - b) <= 1e-10 * np.linalg.norm(b)
        f = F.export(0)
        L, U, rd = f["L"], f["U"], np.asarray(f["rd"])
        assert np.array_equal(np.sort(rd), np.arange(N))
        assert np.allclose(np.diag(L), 1.0) and np.allclose(np.triu(L, 1), 0.0)
        assert np.abs(L).max() <= 1 / (1 - alpha) + 1e-12
        assert np.abs(np.triu(U, 2)).max() == 0 and np.abs(np.tril(U, -2)).max() == 0
        assert (np.abs(np.diag(U, -1)).max() > 0) == blocks
        R = L @ U @ L.T - A[np.ix_(rd, rd)]
        assert np.abs(R).max() <= 1e-12 * np.abs(A).max()
        sgn, ld = np.linalg.slogdet(A)
        ldF = flam_amd.rskelf_logdet(F)
        assert abs(ldF.real - ld) <= 1e-10 * max(1.0, abs(ld))
        assert abs(np.exp(1j * ldF.imag) - sgn) < 1e-10
        afun = (lambda I, J: A[np.ix_(np.asarray(I, dtype=np.int64), np.asarray(J, dtype=np.int64))])
        Fo = ofl.rskelf(afun, pr.x, N, 1e-12, pr.pxyfun, dict(symm="h", order="level"))
        xo = ofl.rskelf_sv(Fo, b)
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("route", ["library", "global"])
def test_large_root_global_paths(route):
    """One box of 2600 points (occ = N): the root's dense LU runs on the
    whole-GPU library path (the default above 384 points) or, with big_n
    raised, in the elimination's global-scratch class; both sweeps run in
    their global workspace slices (rows beyond shared memory); solve and
    apply against the closed form and the exact product."""
    flam_amd, geometry = _dev()
    from flam_amd import tune
    N = 2600
    print(f"large root: circle N={N} occ={N} symm n fp64, {route} elimination", flush=True)
    g = geometry.circle(N)
    saved = tune.get_tuning("big_n")
    try:
        if route == "global":
            tune.set_tuning("big_n", 1 << 30)
        F = flam_amd.rskelf(g.kernel, g.x, N, 1e-12, g.proxy, dict(symm="n"))
    finally:
        tune.set_tuning("big_n", saved)
    assert F.level_table() == [(1, 1, N, 0)]
    pr = opr.CurveDL(N, "circle")
    B = np.random.default_rng(6).standard_normal((N, 3))
    X = flam_amd.rskelf_sv(F, torch.from_numpy(B).cuda()).cpu().numpy()
    Xe = pr.closed_form_inverse_apply(B)
    assert np.linalg.norm(X - Xe) <= 1e-11 * np.linalg.norm(Xe)
    Y = flam_amd.rskelf_mv(F, torch.from_numpy(Xe).cuda()).cpu().numpy()
    assert np.linalg.norm(Y - B) <= 1e-11 * np.linalg.norm(B)
    Xc = flam_amd.rskelf_sv(F, torch.from_numpy(B).cuda(), "c").cpu().numpy()
    assert np.linalg.norm(Xc - Xe) <= 1e-11 * np.linalg.norm(Xe)   # symmetric matrix


@pytest.mark.parametrize("symm", ["n", "p", "h"])
def test_import_export_round_trip(symm):
    """rskx_import of the reference's F layout: the device factor exported
    with rskx_export_factor and imported again gives the same sweeps (the
    inverted diagonal blocks recomputed on the host: agreement to 1e-13)."""
    flam_amd, geometry = _dev()
    N = 4096
    print(f"import/export round trip: circle N={N} symm {symm}", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm=symm))
    Fi = flam_amd.rskelf_import(flam_amd.export_all(F))
    assert Fi.info()["nfactor"] == F.info()["nfactor"]
    B = torch.from_numpy(np.random.default_rng(7).standard_normal((N, 3))).cuda()
    for fn in (flam_amd.rskelf_sv, flam_amd.rskelf_mv):
        for trans in ("n", "c"):
            a, b = fn(F, B, trans), fn(Fi, B, trans)
            assert (a - b).abs().max().item() <= 1e-13 * a.abs().max().item(), (fn, trans)


@pytest.mark.parametrize("kind,symm", [("circle", "p"), ("ellipse", "n")])
def test_oracle_factor_imported_to_the_device(kind, symm):
    """The oracle's factor (the CPU restatement of rskelf.m) imported into
    the device: the device sweeps on it reproduce the oracle's own sweeps
    (sv_*.m, mv_*.m) to 1e-12, independently of the device factorization."""
    flam_amd, geometry = _dev()
    N = 2048
    print(f"oracle factor on the device: {kind} N={N} symm {symm}", flush=True)
    pr = opr.CurveDL(N, kind, a=1.0, b=0.6) if kind == "ellipse" else opr.CurveDL(N, kind)
    Fo = ofl.rskelf(pr.afun, pr.x, 64, 1e-10, pr.pxyfun, dict(symm=symm, order="level"))
    Fd = flam_amd.rskelf_import(dict(Fo, N=N))
    B = np.random.default_rng(8).standard_normal((N, 2))
    for trans in ("n", "c", "t"):
        xs = flam_amd.rskelf_sv(Fd, torch.from_numpy(B).cuda(), trans).cpu().numpy()
        xo = ofl.rskelf_sv(Fo, B, trans)
        assert np.linalg.norm(xs - xo) <= 1e-12 * np.linalg.norm(xo), trans
        ys = flam_amd.rskelf_mv(Fd, torch.from_numpy(B).cuda(), trans).cpu().numpy()
        yo = ofl.rskelf_mv(Fo, B, trans)
        assert np.linalg.norm(ys - yo) <= 1e-12 * np.linalg.norm(yo), trans
