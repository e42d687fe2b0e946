"""Complex factorization (Helmholtz combined field) against the oracle.

The matrix is the exterior Dirichlet combined-field matrix of
oracle/problems.py:CurveCFIE (G = (i/4) H0^(1)(k r), the kernel of
rskelf/test/ie_square3.m), on the star curve r = 1 + 0.3 cos 5t and on
the unit circle (circulant and complex symmetric there).  The device
evaluates H0^(1), H1^(1) with HIP's j0, y0, j1, y1; the oracle with
scipy.special.hankel1, so the product test pins the device kernel to the
library reference.

Tolerances: FP64 products agree with the dense oracle matrix to 1e-12;
per-level skeleton totals within 3% of the oracle (level order); solves
and applies within 100 tol of the dense LU / dense product; FP32 factors
refined by FGMRES to 1e-10.
"""
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


def _np(t):
    return t.detach().cpu().numpy()


def _oracle(g, symm="n"):
    return opr.CurveCFIE.from_arrays(_np(g.x), _np(g.nu), _np(g.w), _np(g.diag), g.k, g.eta,
                                     g.proxy.weight, nproxy=g.proxy.p, symm=symm)


def _cmp_tables(got, ref, frac):
    assert len(got) == len(ref), (got, ref)
    for a, b in zip(got, ref):
        assert (a[0], a[1], a[2]) == (b[0], b[1], b[2]) or (a[0], a[1]) == (b[0], b[1]), (a, b)
        assert abs(a[3] - b[3]) <= max(2, frac * b[3]), (a, b)


def _crand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_cfie_matvec_matches_oracle_dense():
    flam_amd, geometry = _dev()
    N = 1024
    print(f"cfie matvec: star N={N} k=20", flush=True)
    g = geometry.helmholtz_curve(N, "star", k=20.0)
    pr = _oracle(g)
    assert np.abs(_np(g.diag) - opr.CurveCFIE(N, "star", k=20.0).diag).max() < 1e-13
    X = _crand((N, 3), 0)
    Y = _np(flam_amd.kernel_matvec(g.kernel, g.x, torch.from_numpy(X).cuda()))
    Ye = pr.dense() @ X
    assert np.abs(Y - Ye).max() <= 1e-12 * np.abs(Ye).max()


def test_cfie_fp64_star_vs_oracle_and_dense():
    """symm 'n' on the star: rank table against the oracle (two-sided
    compression with the incoming combined-field proxy rows), solves for
    trans 'n', 'c', 't' against dense LU, applies against the dense
    matrix, and the exported factor's structure."""
    flam_amd, geometry = _dev()
    N, occ, tol = 2048, 64, 1e-9
    print(f"cfie fp64: star N={N} k=20 occ={occ} tol={tol} symm n", flush=True)
    g = geometry.helmholtz_curve(N, "star", k=20.0)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="n"))
    pr = _oracle(g)
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="n", order="level"))
    got, ref = F.level_table(), ofl.rank_table(Fo)
    print("device", got, "\noracle", ref, flush=True)
    _cmp_tables(got, ref, 0.03)
    A = pr.dense()
    b = _crand(N, 1)
    for trans, M in (("n", A), ("c", A.conj().T), ("t", A.T)):
        x = _np(flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda(), trans))
        xe = np.linalg.solve(M, b)
        err = np.linalg.norm(x - xe) / np.linalg.norm(xe)
        y = _np(flam_amd.rskelf_mv(F, torch.from_numpy(b).cuda(), trans))
        ye = M @ b
        erry = np.linalg.norm(y - ye) / np.linalg.norm(ye)
        print(f"trans {trans}: solve {err:.2e} apply {erry:.2e}", flush=True)
        assert err < 100 * tol and erry < 100 * tol
    nf = F.info()["nfactor"]
    for i in (0, nf // 2, nf - 1):
        f = F.export(i)
        k, r = len(f["sk"]), len(f["rd"])
        assert f["T"].shape == (k, r) and f["E"].shape == (k, r) and f["F"].shape == (r, k)
        assert np.allclose(np.diag(f["L"]), 1) and np.allclose(np.triu(f["L"], 1), 0)
        assert np.allclose(np.tril(f["U"], -1), 0)
        assert np.array_equal(np.sort(f["p"]), np.arange(r))
    assert len(np.unique(np.concatenate([F.export(i)["rd"] for i in range(nf)]))) == N


def test_cfie_symm_s_on_the_circle():
    """symm 's' (complex symmetric, one-sided compression, T.' updates,
    rskelf.m:220-221; sv_sn / sv_sc / mv_sn / mv_sc): the circle's
    combined-field matrix is complex symmetric; rank table as the
    oracle's 's', solves and applies against the dense matrix, and
    rskelf_sv(.., 't') = rskelf_sv(.., 'n') for the symmetric matrix."""
    flam_amd, geometry = _dev()
    N, occ, tol = 2048, 64, 1e-9
    print(f"cfie fp64: circle N={N} k=20 occ={occ} tol={tol} symm s", flush=True)
    g = geometry.helmholtz_curve(N, "circle", k=20.0)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="s"))
    pr = _oracle(g, symm="s")
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="s", order="level"))
    _cmp_tables(F.level_table(), ofl.rank_table(Fo), 0.03)
    A = pr.dense()
    b = _crand(N, 2)
    for trans, M in (("n", A), ("c", A.conj().T), ("t", A.T)):
        x = _np(flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda(), trans))
        xe = np.linalg.solve(M, b)
        assert np.linalg.norm(x - xe) < 100 * tol * np.linalg.norm(xe), trans
        y = _np(flam_amd.rskelf_mv(F, torch.from_numpy(b).cuda(), trans))
        ye = M @ b
        assert np.linalg.norm(y - ye) < 100 * tol * np.linalg.norm(ye), trans
    xn = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda(), "n")
    xt = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda(), "t")
    assert (xn - xt).abs().max().item() <= 1e-12 * xn.abs().max().item()
    # the circulant product of the device matrix equals the dense oracle
    Yc = _np(g.circulant_matvec(torch.from_numpy(b).cuda()))
    assert np.linalg.norm(Yc - A @ b) <= 1e-11 * np.linalg.norm(A @ b)


def test_cfie_pde_known_answer_through_device_factor():
    """Exterior known answer (a point source inside the star; field outside)
    through the device's FP64 factor, equal to the dense solve's error."""
    flam_amd, geometry = _dev()
    N = 2048
    print(f"cfie pde known answer: star N={N} k=10", flush=True)
    g = geometry.helmholtz_curve(N, "star", k=10.0)
    F = flam_amd.rskelf(g.kernel, g.x, 64, 1e-11, g.proxy, dict(symm="n"))
    pr = _oracle(g)
    A = pr.dense()
    ref = opr.helmholtz_pde_error(pr, lambda B: np.linalg.solve(A, B))
    err = opr.helmholtz_pde_error(pr, lambda B: _np(flam_amd.rskelf_sv(F, torch.from_numpy(B).cuda())))
    print(f"pde error {err:.3e} (dense {ref:.3e})", flush=True)
    assert ref < 1e-6 and abs(err - ref) < 1e-9


@pytest.mark.parametrize("kind,symm", [("star", "n"), ("circle", "s")])
def test_cfie_fp32_refined(kind, symm):
    """Complex FP32 factor (complex64 storage, FP64 ID decisions) as the
    preconditioner of complex FGMRES with the exact FP64 product, to a
    relative residual of 1e-10 on every column."""
    flam_amd, geometry = _dev()
    from flam_amd.refine import fgmres
    N, k = 8192, 40.0
    print(f"cfie fp32 + fgmres: {kind} N={N} k={k} symm {symm}", flush=True)
    g = geometry.helmholtz_curve(N, kind, k=k)
    F = flam_amd.rskelf(g.kernel, g.x, 128, 1e-6, g.proxy, dict(symm=symm, precision="fp32"))
    B = torch.from_numpy(_crand((N, 4), 3)).cuda()
    X, info = fgmres(g.matvec, B, F, tol=1e-10)
    R = B - g.matvec(X)
    rel = (torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(B, dim=0)).max().item()
    print(f"iters {info['iters']} relres {rel:.2e}", flush=True)
    assert rel <= 1e-10
    X1 = flam_amd.rskelf_sv(F, B)
    r1 = (torch.linalg.vector_norm(B - g.matvec(X1), dim=0) /
          torch.linalg.vector_norm(B, dim=0)).max().item()
    print(f"one FP32 solve: relres {r1:.2e}", flush=True)
    assert r1 < 1e-3


def test_cfie_matvec_far_field_against_oracle():
    """At k = 200 most pairs have k r >= 30, where the device evaluates
    H0 and H1 by the Hankel asymptotic expansion (DLMF 10.17.3-4) instead
    of j0 / y0 / j1 / y1: the product against the dense oracle matrix
    (SciPy's hankel1) on the star to 1e-12."""
    flam_amd, geometry = _dev()
    N = 2048
    print(f"cfie matvec far field: star N={N} k=200", flush=True)
    g = geometry.helmholtz_curve(N, "star", k=200.0)
    pr = _oracle(g)
    X = _crand((N, 2), 7)
    Y = _np(flam_amd.kernel_matvec(g.kernel, g.x, torch.from_numpy(X).cuda()))
    Ye = pr.dense() @ X
    assert np.abs(Y - Ye).max() <= 1e-12 * np.abs(Ye).max()
