r"""Large boxes: the wide-box Gram ID and the whole-GPU elimination.

Gram-route boxes beyond the register-resident Gram classes (n > 284) take
the batched wide-box ID: K in global memory, G = K'K on FP64 matrix-core
tiles, a pivoted Cholesky blocked by 32 pivots, T = R11 \ R12 by the Gram
tail or (rank >= 256) by a library triangular solve.  The first test pins
it against the oracle's QR ID on the sphere, whose upper levels hold such
boxes.

Boxes with more than big_n points are eliminated with the whole GPU: a
dense LU / Cholesky (rocSOLVER) with library triangular solves and products
(rskelf.m:218-246).  Lowering big_n sends most non-leaf boxes there.  The
route changes only the order of the arithmetic: the per-level skeleton
totals must agree with the per-workgroup kernels' within 2% (the ranks of
the next level depend on the rounding of M), and the solves must meet the
same accuracy against the FP64 operator.  The sphere case
also takes its root box (no neighbours, everything redundant: a plain dense
LU) through the library LU.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_wide_gram_id_sphere_vs_oracle():
    import flam_amd
    from flam_amd import geometry
    from oracle import flam as ofl
    from oracle import problems as opr
    g = geometry.sphere(5120)
    print("wide Gram ID: sphere nface=5120 occ=128 tol=1e-6 symm=n fp64 id_route=gram", flush=True)
    F = flam_amd.rskelf(g.kernel, g.x, 128, 1e-6, g.proxy, dict(symm="n", id_route="gram"))
    got = F.level_table()
    print("levels:", got, flush=True)
    # the top levels' boxes are beyond the register Gram classes
    assert max(r[2] / r[1] for r in got if r[3] > 0) > 284
    pr = opr.SphereSL.from_arrays(g.x.cpu().numpy(), g.w.cpu().numpy(), g.diag.cpu().numpy(),
                                  g.proxy.pts.cpu().numpy(), symm="n")
    Fo = ofl.rskelf(pr.afun, pr.x, 128, 1e-6, pr.pxyfun, dict(symm="n", order="level"))
    ref = ofl.rank_table(Fo)
    _tables_close(got, ref)
    A = pr.dense()
    b = np.random.default_rng(10).standard_normal(g.N)
    xe = np.linalg.solve(A, b)
    x = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda()).cpu().numpy()
    xo = ofl.rskelf_sv(Fo, b)
    err = np.linalg.norm(x - xe) / np.linalg.norm(xe)
    err_o = np.linalg.norm(xo - xe) / np.linalg.norm(xe)
    print(f"solve error: device {err:.2e}, oracle {err_o:.2e}", flush=True)
    assert err <= max(10 * err_o, 1e-12)


def _pair(flam_amd, tune, big_n, g, occ, tol, opts):
    """Factors with boxes above big_n on the whole-GPU path, and with every
    box on the per-workgroup kernels."""
    saved = tune.get_tuning("big_n")
    try:
        tune.set_tuning("big_n", big_n)
        Fa = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, opts)
        tune.set_tuning("big_n", 1 << 30)
        Fb = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, opts)
    finally:
        tune.set_tuning("big_n", saved)
    return Fa, Fb


def _tables_close(ta, tb):
    assert len(ta) == len(tb)
    for a, b in zip(ta, tb):
        assert (a[0], a[1]) == (b[0], b[1]), (a, b)
        assert abs(a[3] - b[3]) <= max(2, 0.02 * b[3]), (a, b)


def _relres(flam_amd, g, X, B):
    R = B - flam_amd.kernel_matvec(g.kernel, g.x, X)
    return (torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(B, dim=0)).max().item()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_big_path_sphere_lu(prec):
    import flam_amd
    from flam_amd import geometry, tune
    g = geometry.sphere(5120)
    print(f"big path: sphere nface=5120 occ=128 tol=1e-6 symm=n {prec} big_n=200", flush=True)
    opts = dict(symm="n", precision=prec, id_route="gram")
    Fa, Fb = _pair(flam_amd, tune, 200, g, 128, 1e-6, opts)
    ta, tb = Fa.level_table(), Fb.level_table()
    print("big:", ta, "\nper-workgroup:", tb, flush=True)
    _tables_close(ta, tb)
    B = torch.from_numpy(np.random.default_rng(11).standard_normal((g.N, 3))).cuda()
    ea = _relres(flam_amd, g, flam_amd.rskelf_sv(Fa, B), B)
    eb = _relres(flam_amd, g, flam_amd.rskelf_sv(Fb, B), B)
    print(f"relative residual: big {ea:.2e}, per-workgroup {eb:.2e}", flush=True)
    assert ea <= max(3 * eb, 1e-5 if prec == "fp32" else 1e-9)
    # the transposed solve reads the same factor through the other sweeps
    Xt = flam_amd.rskelf_sv(Fa, B, "t")
    Xtb = flam_amd.rskelf_sv(Fb, B, "t")
    dt = ((Xt - Xtb).abs().max() / Xtb.abs().max()).item()
    print(f"transposed solves differ by {dt:.2e}", flush=True)
    assert dt <= (1e-2 if prec == "fp32" else 1e-3)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_big_path_circle_cholesky(prec):
    import flam_amd
    from flam_amd import geometry, tune
    N = 1 << 14
    g = geometry.circle(N)
    print(f"big path: circle N={N} occ=64 tol=1e-6 symm=p {prec} big_n=64", flush=True)
    opts = dict(symm="p", precision=prec, id_route="gram")
    Fa, Fb = _pair(flam_amd, tune, 64, g, 64, 1e-6, opts)
    ta, tb = Fa.level_table(), Fb.level_table()
    _tables_close(ta, tb)
    B = torch.from_numpy(np.random.default_rng(12).standard_normal((N, 2))).cuda()
    Xa, Xb = flam_amd.rskelf_sv(Fa, B), flam_amd.rskelf_sv(Fb, B)
    ea, eb = _relres(flam_amd, g, Xa, B), _relres(flam_amd, g, Xb, B)
    print(f"relative residual: big {ea:.2e}, per-workgroup {eb:.2e}", flush=True)
    assert ea <= max(3 * eb, 1e-5 if prec == "fp32" else 1e-9)
    # the Cholesky factor's log-determinant agrees with the per-workgroup one
    la, lb = flam_amd.rskelf_logdet(Fa), flam_amd.rskelf_logdet(Fb)
    print(f"logdet: big {la}, per-workgroup {lb}", flush=True)
    assert abs(complex(la) - complex(lb)) <= (1e-3 if prec == "fp32" else 1e-6) * max(1.0, abs(complex(lb)))


@pytest.mark.parametrize("symm,prec", [("n", "fp64"), ("n", "fp32"), ("p", "fp64")])
def test_library_solve_matches_kernel_solve(symm, prec):
    """Factors with at least lib_solve_r redundant points are solved by
    library triangular solves and products; forced down to 64, most
    non-leaf factors take that path, and the solves (both transposes for
    'n') must agree with the per-workgroup sweeps to rounding."""
    import flam_amd
    from flam_amd import geometry, tune
    if symm == "n":
        g = geometry.sphere(5120)
        occ, tol = 128, 1e-6
    else:
        g = geometry.circle(1 << 14)
        occ, tol = 64, 1e-8
    print(f"library solve: {g.N} points symm={symm} {prec} lib_solve_r=64", flush=True)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm=symm, precision=prec))
    B = torch.from_numpy(np.random.default_rng(13).standard_normal((g.N, 5))).cuda()
    rtol = 1e-4 if prec == "fp32" else 1e-10
    saved = tune.get_tuning("lib_solve_r")
    for trans in (("n", "t") if symm == "n" else ("n",)):
        tune.set_tuning("lib_solve_r", 0)
        try:
            Xk = flam_amd.rskelf_sv(F, B, trans)
            tune.set_tuning("lib_solve_r", 64)
            Xl = flam_amd.rskelf_sv(F, B, trans)
        finally:
            tune.set_tuning("lib_solve_r", saved)
        d = ((Xl - Xk).abs().max() / Xk.abs().max()).item()
        print(f"trans {trans}: max relative difference {d:.2e}", flush=True)
        assert d <= rtol
